// sx_mlp.hip -- the learned STOKE proposal: a two-layer MLP over the target's
// instructions, trained and run on the device (sx_mlp_* in include/sx6502.h).
//
//   tokens (8 per example: Go-order instruction index, 2863 = padding)
//   -> embeddings E[2864][64] (bf16), concatenated to X[B][512]
//   -> H = relu(X W1 + b1), W1[512][256]
//   -> Z = H W2 + b2, W2[256][2871]  -> softmax over 2,871 outputs: the 2,863
//      instructions, then 8 pointers ("the instruction at context position j").
// A target's proposal folds each pointer's mass onto the instruction at that
// position (sx_mlp_proposal).
//
// Weights are used in bf16 with fp32 master copies for Adam.  All products
// run on MFMA (gfx950's v_mfma_f32_16x16x32_bf16: bf16 inputs, fp32
// accumulation, 32 of K per instruction) in one tiled kernel with generic
// strides, so the transposed products of the backward pass need no copies.  The training step (batch draw, forward,
// loss, backward, Adam) is captured once in a hipGraph and replayed; its
// step counter lives in device memory.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/sx6502.h"

hipStream_t sx_internal_stream(sx_handle *h);
int sx_internal_error(sx_handle *h, int code, const char *msg);

namespace {

constexpr int kTok = 2864;   // 2,863 instructions + padding
constexpr int kInsts = 2863;  // instructions
constexpr int kCtx = 8;
constexpr int kEmb = 64;
constexpr int kIn = kCtx * kEmb;  // 512
constexpr int kHid = 256;
constexpr int kOut = kInsts + 8;  // outputs: 2,863 instructions, then 8 context positions
constexpr int kOutPad = 2880;     // outputs padded to a multiple of 64

constexpr size_t kOffE = 0;
constexpr size_t kOffW1 = kOffE + (size_t)kTok * kEmb;
constexpr size_t kOffB1 = kOffW1 + (size_t)kIn * kHid;
constexpr size_t kOffW2 = kOffB1 + kHid;
constexpr size_t kOffB2 = kOffW2 + (size_t)kHid * kOut;
constexpr size_t kParams = kOffB2 + kOut;

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// C[M][N] (fp32, ldc) = sum_k A(m,k) B(k,n) with A(m,k) = A[m*sam + k*sak] and
// B(k,n) = B[k*sbk + n*sbn] (bf16).  64x64 tile per 256-thread block, each
// wave a 32x32 quarter as 2x2 MFMA 16x16 blocks; K in steps of 32 staged in
// LDS, one MFMA k-step of 32 per stage (lane l holds A rows l % 16 and B
// columns l % 16 at k = 8 (l / 16) + j, j < 8: the same k for both operands,
// so the products pair up whatever order the unit sums them in).  gridDim.z > 1 splits K: block z
// takes K range [z * kchunk, ...) and adds its partial product to C with
// atomics (C zeroed by the caller; no epilogue).  Otherwise the epilogue
// applies + bias[n] (if bias), relu (if relu), + C (if accumulate) and
// writes an optional bf16 copy Cb.
__global__ __launch_bounds__(256) void k_gemm(const uint16_t *__restrict__ A, long sam, long sak,
                                              const uint16_t *__restrict__ B, long sbk, long sbn, float *C,
                                              long ldc, uint16_t *Cb, int M, int N, int K,
                                              const float *__restrict__ bias, int relu, int accumulate) {
  __shared__ uint16_t As[64][32 + 4];
  __shared__ uint16_t Bs[32][64 + 4];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;
  const int kchunk = (((K + gridDim.z - 1) / gridDim.z) + 31) & ~31;
  const int kbeg = blockIdx.z * kchunk, kend = min(K, kbeg + kchunk);
  v4f acc[2][2];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
  for (int k0 = kbeg; k0 < kend; k0 += 32) {
    // stage A: 64 x 32, B: 32 x 64; consecutive threads walk the operand's
    // contiguous dimension so the loads coalesce
    if (sak == 1) {
      for (int e = tid; e < 64 * 32; e += 256) {
        const int r = e >> 5, k = e & 31, gm = m0 + r, gk = k0 + k;
        As[r][k] = (gm < M && gk < kend) ? A[(long)gm * sam + gk] : (uint16_t)0;
      }
    } else {
      for (int e = tid; e < 64 * 32; e += 256) {
        const int r = e & 63, k = e >> 6, gm = m0 + r, gk = k0 + k;
        As[r][k] = (gm < M && gk < kend) ? A[(long)gm * sam + (long)gk * sak] : (uint16_t)0;
      }
    }
    if (sbn == 1) {
      for (int e = tid; e < 32 * 64; e += 256) {
        const int k = e >> 6, c = e & 63, gk = k0 + k, gn = n0 + c;
        Bs[k][c] = (gk < kend && gn < N) ? B[(long)gk * sbk + gn] : (uint16_t)0;
      }
    } else {
      for (int e = tid; e < 32 * 64; e += 256) {
        const int k = e & 31, c = e >> 5, gk = k0 + k, gn = n0 + c;
        Bs[k][c] = (gk < kend && gn < N) ? B[(long)gk * sbk + (long)gn * sbn] : (uint16_t)0;
      }
    }
    __syncthreads();
    const int l16 = lane & 15, kq = (lane >> 4) * 8;
    v8bf a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int q = 0; q < 8; q++) a[i][q] = __builtin_bit_cast(__bf16, As[wm + i * 16 + l16][kq + q]);
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int q = 0; q < 8; q++) b[j][q] = __builtin_bit_cast(__bf16, Bs[kq + q][wn + j * 16 + l16]);
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  const bool split = gridDim.z > 1;
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++)
      for (int q = 0; q < 4; q++) {
        int r = m0 + wm + i * 16 + (lane >> 4) * 4 + q;
        int c = n0 + wn + j * 16 + (lane & 15);
        if (r < M && c < N) {
          float v = acc[i][j][q];
          if (split) {
            atomicAdd(&C[(long)r * ldc + c], v);
            continue;
          }
          if (accumulate) v += C[(long)r * ldc + c];
          if (bias) v += bias[c];
          if (relu) v = v > 0.f ? v : 0.f;
          C[(long)r * ldc + c] = v;
          if (Cb) Cb[(long)r * ldc + c] = f2bf(v);
        }
      }
}

__device__ __forceinline__ void philox_d(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
                                         uint32_t out[4]) {
  for (int r = 0; r < 10; r++) {
    uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
    uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
    uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Batch row b: example index (drawn with Philox from the step counter when
// sample != 0, else row b of the given set); gathers the embeddings into X.
__global__ void k_gather(const uint16_t *ctx_all, const uint16_t *lab_all, uint32_t n_all, int sample,
                         const uint32_t *step_ctr, uint32_t k0, uint32_t k1, const uint16_t *Eb, uint16_t *X,
                         uint16_t *tok, uint16_t *lab, int B) {
  int b = blockIdx.x;
  if (b >= B) return;
  __shared__ uint32_t ex;
  if (threadIdx.x == 0) {
    uint32_t idx = (uint32_t)b;
    if (sample) {
      uint32_t r[4];
      philox_d(*step_ctr, (uint32_t)b, 0x5354u, 0, k0, k1, r);
      idx = __umulhi(r[0], n_all);
    }
    ex = idx;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kIn; e += blockDim.x) {
    int s = e / kEmb, d = e % kEmb;
    uint16_t t = ctx_all[(size_t)ex * kCtx + s];
    X[(size_t)b * kIn + e] = Eb[(size_t)t * kEmb + d];
    if (d == 0) tok[b * kCtx + s] = t;
  }
  if (threadIdx.x == 0) lab[b] = lab_all[ex];
}

// Softmax cross-entropy per row; dZ = (p - onehot) / B in fp32 and bf16.
__global__ void k_softmax_ce(const float *Z, long ldz, const uint16_t *lab, int B, float *dZ, uint16_t *dZb,
                             float *loss_rows, float *probs) {
  int b = blockIdx.x;
  __shared__ float red[256];
  const float *z = Z + (long)b * ldz;
  float mx = -INFINITY;
  for (int c = threadIdx.x; c < kOut; c += blockDim.x) mx = fmaxf(mx, z[c]);
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  mx = red[0];
  __syncthreads();
  float sum = 0.f;
  for (int c = threadIdx.x; c < kOut; c += blockDim.x) sum += expf(z[c] - mx);
  red[threadIdx.x] = sum;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  sum = red[0];
  const float inv = 1.f / sum;
  const int y = lab ? lab[b] : -1;
  for (int c = threadIdx.x; c < kOutPad; c += blockDim.x) {
    float p = c < kOut ? expf(z[c] - mx) * inv : 0.f;
    if (probs && c < kOut) probs[(long)b * kOut + c] = p;
    if (dZ) {
      float g = c < kOut ? (p - (c == y ? 1.f : 0.f)) / (float)B : 0.f;
      dZ[(long)b * kOutPad + c] = g;
      dZb[(long)b * kOutPad + c] = f2bf(g);
    }
  }
  if (threadIdx.x == 0 && loss_rows && y >= 0) loss_rows[b] = -(z[y] - mx - logf(sum));
}

// column sums: out[n] += sum_b in[b][n] (out zeroed by the caller); block
// (x, y) takes 64 columns and rows [y kColRows, (y + 1) kColRows), its 4 waves
// a quarter of those each, reduced through LDS and added with one atomic per
// column (a block per 64 columns alone left the chip mostly idle).
constexpr int kColRows = 128;
__global__ __launch_bounds__(256) void k_colsum(const float *in, long ld, int B, int N, float *out) {
  __shared__ float part[4][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int n = blockIdx.x * 64 + lane;
  const int b1 = min(B, (int)(blockIdx.y + 1) * kColRows);
  float acc = 0.f;
  if (n < N)
    for (int b = (int)blockIdx.y * kColRows + wv; b < b1; b += 4) acc += in[(long)b * ld + n];
  part[wv][lane] = acc;
  __syncthreads();
  if (wv == 0 && n < N) atomicAdd(&out[n], part[0][lane] + part[1][lane] + part[2][lane] + part[3][lane]);
}

// dH *= relu'(H), and a bf16 copy
__global__ void k_relu_back(float *dH, const float *H, uint16_t *dHb, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float g = H[i] > 0.f ? dH[i] : 0.f;
  dH[i] = g;
  dHb[i] = f2bf(g);
}

// The embedding gradient dE[t] = sum of the dX rows (example b, context
// position s) whose token is t, as segmented sums instead of one atomic per
// element (the padding token alone fills most contexts, and its atomics on
// one row of dE serialised).  k_emb_sort (one block) lists the batch's
// positions b * 8 + s grouped by token: off[t] .. off[t + 1] in pos.  (The
// padding token kTok - 1 is a token like the others.)
__global__ __launch_bounds__(1024) void k_emb_sort(const uint16_t *tok, int B, uint32_t *off, uint32_t *pos) {
  __shared__ uint32_t cnt[kTok + 1];
  __shared__ uint32_t wsum[16];
  const int n = B * kCtx, tid = threadIdx.x;
  for (int t = tid; t <= kTok; t += 1024) cnt[t] = 0;
  __syncthreads();
  for (int i = tid; i < n; i += 1024) atomicAdd(&cnt[tok[i]], 1u);
  __syncthreads();
  // exclusive scan of cnt[0 .. kTok] in place: 3 entries per thread
  constexpr int kPer = (kTok + 1 + 1023) / 1024;
  uint32_t loc[kPer], run = 0;
#pragma unroll
  for (int k = 0; k < kPer; k++) {
    const int t = tid * kPer + k;
    loc[k] = t <= kTok ? cnt[t] : 0u;
    run += loc[k];
  }
  // inclusive scan of the per-thread totals across the block
  uint32_t x = run;
  const int lane = tid & 63, w = tid >> 6;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  uint32_t before = 0;
  for (int k = 0; k < w; k++) before += wsum[k];
  uint32_t acc = before + x - run;  // exclusive prefix of this thread's first entry
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kPer; k++) {
    const int t = tid * kPer + k;
    if (t <= kTok) {
      cnt[t] = acc;
      off[t] = acc;
    }
    acc += loc[k];
  }
  __syncthreads();
  for (int i = tid; i < n; i += 1024) pos[atomicAdd(&cnt[tok[i]], 1u)] = (uint32_t)i;
}
// The sums: a wave takes 16 consecutive listed positions (positions of one
// token are consecutive), lane d the embedding dimension d, and adds its
// running sum to dE (zeroed with the other gradients) when the token changes
// and at the end: a token's rows spread over many waves, with one atomic per
// wave and dimension where a block per token left the padding token's
// thousand rows to one block.
constexpr int kEmbPerWave = 16;
__global__ __launch_bounds__(256) void k_emb_grad(const float *dX, const uint16_t *tok, const uint32_t *pos, int n,
                                                  float *dE) {
  const int lane = threadIdx.x & 63;
  const int i0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * kEmbPerWave;
  if (i0 >= n) return;
  const int i1 = min(n, i0 + kEmbPerWave);
  uint32_t cur = kTok;
  float acc = 0.f;
  for (int i = i0; i < i1; i++) {
    const uint32_t q = pos[i];  // b * kCtx + s
    const uint32_t t = tok[q];
    if (t != cur) {
      if (cur != kTok) atomicAdd(&dE[(size_t)cur * kEmb + lane], acc);
      cur = t;
      acc = 0.f;
    }
    acc += dX[(size_t)(q / kCtx) * kIn + (q % kCtx) * kEmb + lane];
  }
  atomicAdd(&dE[(size_t)cur * kEmb + lane], acc);
}

__global__ void k_fill_f32(float *p, size_t n, float v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Adam with bias correction; t from the device step counter (incremented by
// k_step_inc after the update).  Writes the bf16 working copy.
__global__ void k_adam(float *p, const float *g, float *m, float *v, uint16_t *pb, size_t n, const uint32_t *step_ctr,
                       float lr, float b1, float b2, float eps) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float t = (float)(*step_ctr + 1);
  float gi = g[i];
  float mi = b1 * m[i] + (1.f - b1) * gi;
  float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  float mh = mi / (1.f - powf(b1, t)), vh = vi / (1.f - powf(b2, t));
  float pi = p[i] - lr * mh / (sqrtf(vh) + eps);
  p[i] = pi;
  pb[i] = f2bf(pi);
}

__global__ void k_step_inc(uint32_t *step_ctr) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *step_ctr += 1;
}

__global__ void k_to_bf16(const float *p, uint16_t *pb, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) pb[i] = f2bf(p[i]);
}

// mean of the B row losses, one 256-thread block
__global__ __launch_bounds__(256) void k_sum_loss(const float *rows, int B, float *out) {
  __shared__ float red[256];
  float s = 0.f;
  for (int b = threadIdx.x; b < B; b += 256) s += rows[b];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0] / (float)B;
}

}  // namespace

struct sx_mlp {
  sx_handle *h = nullptr;
  int B = 0;  // batch capacity of the work buffers
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  uint16_t *pb = nullptr;
  // work buffers
  uint16_t *X = nullptr, *tok = nullptr, *lab = nullptr, *Hb = nullptr, *dZb = nullptr, *dHb = nullptr;
  float *H = nullptr, *Z = nullptr, *dZ = nullptr, *dH = nullptr, *dX = nullptr, *lossr = nullptr,
        *loss = nullptr;
  uint32_t *step = nullptr;
  uint32_t *eoff = nullptr, *epos = nullptr;  // the batch's positions grouped by token (embedding gradient)
  uint16_t *ctx_d = nullptr, *lab_d = nullptr;
  size_t n_data = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  float graph_lr = 0.f;
  uint64_t graph_seed = 0;
  uint32_t graph_n = 0;
};

namespace {

inline unsigned nb(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

int alloc_work(sx_mlp *m, int B) {
  if (m->B >= B) return SX_OK;
  for (void *q : {(void *)m->X, (void *)m->tok, (void *)m->lab, (void *)m->Hb, (void *)m->dZb, (void *)m->dHb,
                  (void *)m->H, (void *)m->Z, (void *)m->dZ, (void *)m->dH, (void *)m->dX, (void *)m->lossr,
                  (void *)m->eoff, (void *)m->epos})
    if (q) (void)hipFree(q);
  m->X = m->tok = m->lab = m->Hb = m->dZb = m->dHb = nullptr;
  m->H = m->Z = m->dZ = m->dH = m->dX = m->lossr = nullptr;
  m->eoff = m->epos = nullptr;
  m->B = 0;
  if (m->gexec) { (void)hipGraphExecDestroy(m->gexec); m->gexec = nullptr; }
  if (m->graph) { (void)hipGraphDestroy(m->graph); m->graph = nullptr; }
  bool ok = hipMalloc(&m->X, (size_t)B * kIn * 2) == hipSuccess && hipMalloc(&m->tok, (size_t)B * kCtx * 2) == hipSuccess &&
            hipMalloc(&m->lab, (size_t)B * 2) == hipSuccess && hipMalloc(&m->Hb, (size_t)B * kHid * 2) == hipSuccess &&
            hipMalloc(&m->dZb, (size_t)B * kOutPad * 2) == hipSuccess &&
            hipMalloc(&m->dHb, (size_t)B * kHid * 2) == hipSuccess && hipMalloc(&m->H, (size_t)B * kHid * 4) == hipSuccess &&
            hipMalloc(&m->Z, (size_t)B * kOutPad * 4) == hipSuccess && hipMalloc(&m->dZ, (size_t)B * kOutPad * 4) == hipSuccess &&
            hipMalloc(&m->dH, (size_t)B * kHid * 4) == hipSuccess && hipMalloc(&m->dX, (size_t)B * kIn * 4) == hipSuccess &&
            hipMalloc(&m->lossr, (size_t)B * 4) == hipSuccess &&
            hipMalloc(&m->eoff, (size_t)(kTok + 1) * 4) == hipSuccess &&
            hipMalloc(&m->epos, (size_t)B * kCtx * 4) == hipSuccess;
  if (!ok) return sx_internal_error(m->h, SX_ENOMEM, "mlp work buffers");
  m->B = B;
  return SX_OK;
}

// forward (+ loss and backward into m->g when train) for rows already gathered into X/tok/lab.
void enqueue_fwd_bwd(sx_mlp *m, int B, bool backward, float *probs, hipStream_t s) {
  const uint16_t *Eb = m->pb + kOffE, *W1b = m->pb + kOffW1, *W2b = m->pb + kOffW2;
  // H = relu(X W1 + b1)
  hipLaunchKernelGGL(k_gemm, dim3(nb(kHid, 64), nb(B, 64)), dim3(256), 0, s, m->X, (long)kIn, 1L, W1b, (long)kHid, 1L,
                     m->H, (long)kHid, m->Hb, B, kHid, kIn, m->p + kOffB1, 1, 0);
  // Z = H W2 + b2
  hipLaunchKernelGGL(k_gemm, dim3(nb(kOut, 64), nb(B, 64)), dim3(256), 0, s, m->Hb, (long)kHid, 1L, W2b, (long)kOut,
                     1L, m->Z, (long)kOutPad, (uint16_t *)nullptr, B, kOut, kHid, m->p + kOffB2, 0, 0);
  hipLaunchKernelGGL(k_softmax_ce, dim3(B), dim3(256), 0, s, m->Z, (long)kOutPad, backward ? m->lab : nullptr, B,
                     backward ? m->dZ : nullptr, backward ? m->dZb : nullptr, backward ? m->lossr : nullptr, probs);
  if (!backward) return;
  hipLaunchKernelGGL(k_sum_loss, dim3(1), dim3(256), 0, s, m->lossr, B, m->loss);
  hipLaunchKernelGGL(k_fill_f32, dim3(nb(kParams, 256)), dim3(256), 0, s, m->g, kParams, 0.f);
  // products over the batch (K = B) and over the 2,871 outputs are split in K
  // and accumulated with atomics into zeroed outputs: the output tiles alone
  // are too few to fill the chip
  const unsigned split_b = (unsigned)std::max(1, B / 256);
  // dW2 = H^T dZ   (A = Hb viewed transposed: A(m=k_hid, k=b) = Hb[b*kHid + m])
  hipLaunchKernelGGL(k_gemm, dim3(nb(kOut, 64), nb(kHid, 64), split_b), dim3(256), 0, s, m->Hb, 1L, (long)kHid, m->dZb,
                     (long)kOutPad, 1L, m->g + kOffW2, (long)kOut, (uint16_t *)nullptr, kHid, kOut, B,
                     (const float *)nullptr, 0, 0);
  hipLaunchKernelGGL(k_colsum, dim3(nb(kOut, 64), nb(B, kColRows)), dim3(256), 0, s, m->dZ, (long)kOutPad, B, kOut,
                     m->g + kOffB2);
  // dH = dZ W2^T  (B(k=c, n=j) = W2[j*kOut + c])
  hipLaunchKernelGGL(k_fill_f32, dim3(nb((size_t)B * kHid, 256)), dim3(256), 0, s, m->dH, (size_t)B * kHid, 0.f);
  hipLaunchKernelGGL(k_gemm, dim3(nb(kHid, 64), nb(B, 64), 8), dim3(256), 0, s, m->dZb, (long)kOutPad, 1L, W2b, 1L,
                     (long)kOut, m->dH, (long)kHid, (uint16_t *)nullptr, B, kHid, kOut, (const float *)nullptr, 0, 0);
  hipLaunchKernelGGL(k_relu_back, dim3(nb((size_t)B * kHid, 256)), dim3(256), 0, s, m->dH, m->H, m->dHb, B * kHid);
  hipLaunchKernelGGL(k_colsum, dim3(nb(kHid, 64), nb(B, kColRows)), dim3(256), 0, s, m->dH, (long)kHid, B, kHid,
                     m->g + kOffB1);
  // dW1 = X^T dH
  hipLaunchKernelGGL(k_gemm, dim3(nb(kHid, 64), nb(kIn, 64), (unsigned)std::max(1, B / 128)), dim3(256), 0, s, m->X, 1L,
                     (long)kIn, m->dHb, (long)kHid,
                     1L, m->g + kOffW1, (long)kHid, (uint16_t *)nullptr, kIn, kHid, B, (const float *)nullptr, 0, 0);
  // dX = dH W1^T, scattered into dE
  hipLaunchKernelGGL(k_gemm, dim3(nb(kIn, 64), nb(B, 64)), dim3(256), 0, s, m->dHb, (long)kHid, 1L, W1b, 1L,
                     (long)kHid, m->dX, (long)kIn, (uint16_t *)nullptr, B, kIn, kHid, (const float *)nullptr, 0, 0);
  hipLaunchKernelGGL(k_emb_sort, dim3(1), dim3(1024), 0, s, m->tok, B, m->eoff, m->epos);
  hipLaunchKernelGGL(k_emb_grad, dim3(nb((size_t)B * kCtx, 4 * kEmbPerWave)), dim3(256), 0, s, m->dX, m->tok,
                     m->epos, B * kCtx, m->g + kOffE);
}

int upload_set(sx_mlp *m, const uint16_t *ctx, const uint16_t *labels, size_t n) {
  (void)hipFree(m->ctx_d);
  (void)hipFree(m->lab_d);
  m->ctx_d = m->lab_d = nullptr;
  if (hipMalloc(&m->ctx_d, n * kCtx * 2 + 16) != hipSuccess || hipMalloc(&m->lab_d, n * 2 + 16) != hipSuccess)
    return sx_internal_error(m->h, SX_ENOMEM, "mlp data");
  hipStream_t s = sx_internal_stream(m->h);
  if (hipMemcpyAsync(m->ctx_d, ctx, n * kCtx * 2, hipMemcpyHostToDevice, s) != hipSuccess) return SX_EHIP;
  if (labels && hipMemcpyAsync(m->lab_d, labels, n * 2, hipMemcpyHostToDevice, s) != hipSuccess) return SX_EHIP;
  if (!labels && hipMemsetAsync(m->lab_d, 0, n * 2, s) != hipSuccess) return SX_EHIP;
  m->n_data = n;
  return hipStreamSynchronize(s) == hipSuccess ? SX_OK : SX_EHIP;
}

int check_tokens(sx_mlp *m, const uint16_t *ctx, const uint16_t *labels, size_t n) {
  for (size_t i = 0; i < n * kCtx; i++)
    if (ctx[i] >= kTok) return sx_internal_error(m->h, SX_EINVAL, "context token out of range");
  if (labels)
    for (size_t i = 0; i < n; i++)
      if (labels[i] >= kOut) return sx_internal_error(m->h, SX_EINVAL, "label out of range");
  return SX_OK;
}

}  // namespace

extern "C" uint64_t sx_mlp_param_count(void) { return kParams; }

extern "C" int sx_mlp_create(sx_handle *h, uint64_t seed, sx_mlp **out) {
  if (!h || !out) return SX_EINVAL;
  sx_mlp *m = new sx_mlp();
  m->h = h;
  bool ok = hipMalloc(&m->p, kParams * 4) == hipSuccess && hipMalloc(&m->g, kParams * 4) == hipSuccess &&
            hipMalloc(&m->m, kParams * 4) == hipSuccess && hipMalloc(&m->v, kParams * 4) == hipSuccess &&
            hipMalloc(&m->pb, kParams * 2) == hipSuccess && hipMalloc(&m->step, 16) == hipSuccess &&
            hipMalloc(&m->loss, 16) == hipSuccess;
  if (!ok) {
    sx_mlp_destroy(m);
    return sx_internal_error(h, SX_ENOMEM, "mlp parameters");
  }
  // initial weights: uniform(-a, a) with a = 1/sqrt(fan_in) (embeddings: 0.1), zero biases
  std::vector<float> p(kParams, 0.f);
  uint64_t st = seed * 0x9E3779B97F4A7C15ull + 1;
  auto rnd = [&]() {
    st ^= st << 13; st ^= st >> 7; st ^= st << 17;  // xorshift64
    return (float)((st >> 40) * (1.0 / 16777216.0)) * 2.f - 1.f;
  };
  for (size_t i = kOffE; i < kOffW1; i++) p[i] = 0.1f * rnd();
  for (size_t i = kOffW1; i < kOffB1; i++) p[i] = rnd() / sqrtf((float)kIn);
  for (size_t i = kOffW2; i < kOffB2; i++) p[i] = rnd() / sqrtf((float)kHid);
  int rc = sx_mlp_set_params(m, p.data());
  if (rc) {
    sx_mlp_destroy(m);
    return rc;
  }
  *out = m;
  return SX_OK;
}

extern "C" int sx_mlp_destroy(sx_mlp *m) {
  if (!m) return SX_EINVAL;
  if (m->gexec) (void)hipGraphExecDestroy(m->gexec);
  if (m->graph) (void)hipGraphDestroy(m->graph);
  for (void *q : {(void *)m->p, (void *)m->g, (void *)m->m, (void *)m->v, (void *)m->pb, (void *)m->step,
                  (void *)m->loss, (void *)m->X, (void *)m->tok, (void *)m->lab, (void *)m->Hb, (void *)m->dZb,
                  (void *)m->dHb, (void *)m->H, (void *)m->Z, (void *)m->dZ, (void *)m->dH, (void *)m->dX,
                  (void *)m->lossr, (void *)m->ctx_d, (void *)m->lab_d, (void *)m->eoff, (void *)m->epos})
    if (q) (void)hipFree(q);
  delete m;
  return SX_OK;
}

extern "C" int sx_mlp_get_params(sx_mlp *m, float *out) {
  if (!m || !out) return SX_EINVAL;
  hipStream_t s = sx_internal_stream(m->h);
  if (hipMemcpyAsync(out, m->p, kParams * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return sx_internal_error(m->h, SX_EHIP, "mlp get params");
  return SX_OK;
}

extern "C" int sx_mlp_set_params(sx_mlp *m, const float *in) {
  if (!m || !in) return SX_EINVAL;
  hipStream_t s = sx_internal_stream(m->h);
  bool ok = hipMemcpyAsync(m->p, in, kParams * 4, hipMemcpyHostToDevice, s) == hipSuccess &&
            hipMemsetAsync(m->m, 0, kParams * 4, s) == hipSuccess && hipMemsetAsync(m->v, 0, kParams * 4, s) == hipSuccess &&
            hipMemsetAsync(m->step, 0, 16, s) == hipSuccess;
  if (ok) {
    hipLaunchKernelGGL(k_to_bf16, dim3(nb(kParams, 256)), dim3(256), 0, s, m->p, m->pb, kParams);
    ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(s) == hipSuccess;
  }
  return ok ? SX_OK : sx_internal_error(m->h, SX_EHIP, "mlp set params");
}

extern "C" int sx_mlp_loss_grad(sx_mlp *m, const uint16_t *ctx, const uint16_t *labels, uint32_t n, float *loss,
                                float *grad) {
  if (!m || !ctx || !labels || n == 0 || !loss || !grad) return SX_EINVAL;
  int rc = check_tokens(m, ctx, labels, n);
  if (rc) return rc;
  if ((rc = alloc_work(m, (int)n)) || (rc = upload_set(m, ctx, labels, n))) return rc;
  hipStream_t s = sx_internal_stream(m->h);
  hipLaunchKernelGGL(k_gather, dim3(n), dim3(256), 0, s, m->ctx_d, m->lab_d, (uint32_t)n, 0, m->step, 0u, 0u,
                     m->pb + kOffE, m->X, m->tok, m->lab, (int)n);
  enqueue_fwd_bwd(m, (int)n, true, nullptr, s);
  if (hipGetLastError() != hipSuccess) return sx_internal_error(m->h, SX_EHIP, "mlp launch");
  if (hipMemcpyAsync(grad, m->g, kParams * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(loss, m->loss, 4, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
    return sx_internal_error(m->h, SX_EHIP, "mlp loss/grad");
  return SX_OK;
}

extern "C" int sx_mlp_train(sx_mlp *m, const uint16_t *ctx, const uint16_t *labels, uint32_t n, uint32_t steps,
                            uint32_t batch, float lr, uint64_t seed, float *losses) {
  if (!m || !ctx || !labels || n == 0 || batch == 0 || batch > 4096) return SX_EINVAL;
  int rc = check_tokens(m, ctx, labels, n);
  if (rc) return rc;
  if ((rc = alloc_work(m, (int)batch)) || (rc = upload_set(m, ctx, labels, n))) return rc;
  hipStream_t s = sx_internal_stream(m->h);
  // capture one training step into a graph (kernels only: no memset nodes)
  if (m->gexec) {
    (void)hipGraphExecDestroy(m->gexec);
    m->gexec = nullptr;
  }
  if (m->graph) {
    (void)hipGraphDestroy(m->graph);
    m->graph = nullptr;
  }
  if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return sx_internal_error(m->h, SX_EHIP, "graph capture");
  const int B = (int)batch;
  hipLaunchKernelGGL(k_gather, dim3(B), dim3(256), 0, s, m->ctx_d, m->lab_d, n, 1, m->step, (uint32_t)seed,
                     (uint32_t)(seed >> 32), m->pb + kOffE, m->X, m->tok, m->lab, B);
  enqueue_fwd_bwd(m, B, true, nullptr, s);
  hipLaunchKernelGGL(k_adam, dim3(nb(kParams, 256)), dim3(256), 0, s, m->p, m->g, m->m, m->v, m->pb, kParams, m->step,
                     lr, 0.9f, 0.999f, 1e-8f);
  hipLaunchKernelGGL(k_step_inc, dim3(1), dim3(64), 0, s, m->step);
  if (hipStreamEndCapture(s, &m->graph) != hipSuccess ||
      hipGraphInstantiate(&m->gexec, m->graph, nullptr, nullptr, 0) != hipSuccess)
    return sx_internal_error(m->h, SX_EHIP, "graph instantiate");
  for (uint32_t it = 0; it < steps; it++) {
    if (hipGraphLaunch(m->gexec, s) != hipSuccess) return sx_internal_error(m->h, SX_EHIP, "graph launch");
    if (losses && hipMemcpyAsync(&losses[it], m->loss, 4, hipMemcpyDeviceToHost, s) != hipSuccess)
      return sx_internal_error(m->h, SX_EHIP, "loss copy");
    if ((it & 255) == 255 && hipStreamSynchronize(s) != hipSuccess) return sx_internal_error(m->h, SX_EHIP, "train");
  }
  if (hipStreamSynchronize(s) != hipSuccess) return sx_internal_error(m->h, SX_EHIP, "train");
  return SX_OK;
}

extern "C" int sx_mlp_predict(sx_mlp *m, const uint16_t *ctx, uint32_t n, float *probs) {
  if (!m || !ctx || n == 0 || !probs) return SX_EINVAL;
  int rc = check_tokens(m, ctx, nullptr, n);
  if (rc) return rc;
  if ((rc = alloc_work(m, (int)std::min<uint32_t>(n, 1024))) ) return rc;
  hipStream_t s = sx_internal_stream(m->h);
  float *pd = nullptr;
  if (hipMalloc(&pd, (size_t)std::min<uint32_t>(n, 1024) * kOut * 4) != hipSuccess)
    return sx_internal_error(m->h, SX_ENOMEM, "predict buffer");
  if ((rc = upload_set(m, ctx, nullptr, n))) {
    (void)hipFree(pd);
    return rc;
  }
  for (uint32_t b0 = 0; b0 < n; b0 += 1024) {
    int B = (int)std::min<uint32_t>(1024, n - b0);
    hipLaunchKernelGGL(k_gather, dim3(B), dim3(256), 0, s, m->ctx_d + (size_t)b0 * kCtx, m->lab_d + b0, (uint32_t)B, 0,
                       m->step, 0u, 0u, m->pb + kOffE, m->X, m->tok, m->lab, B);
    enqueue_fwd_bwd(m, B, false, pd, s);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(probs + (size_t)b0 * kOut, pd, (size_t)B * kOut * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      (void)hipFree(pd);
      return sx_internal_error(m->h, SX_EHIP, "predict");
    }
  }
  (void)hipFree(pd);
  return SX_OK;
}

extern "C" int sx_proposal_weights(const float *probs, uint64_t n, double eps, uint32_t *out) {
  if (!probs || !out || !(eps > 0.0) || eps > 1.0) return SX_EINVAL;
  // floor(w * (2^32 - 2863)) + 1 per instruction sums to at most 2^32; the
  // remainder goes to the largest weight.
  const double scale = 4294967296.0 - (double)kInsts;
  for (uint64_t r = 0; r < n; r++) {
    const float *p = probs + r * kInsts;
    uint32_t *o = out + r * kInsts;
    double tot = 0.0;
    for (int i = 0; i < kInsts; i++) {
      if (!(p[i] >= 0.f)) return SX_EINVAL;
      tot += p[i];
    }
    if (!(tot > 0.0)) return SX_EINVAL;
    uint64_t sum = 0;
    int best = 0;
    for (int i = 0; i < kInsts; i++) {
      double w = (1.0 - eps) * (double)p[i] / tot + eps / (double)kInsts;
      uint64_t u = (uint64_t)std::floor(w * scale) + 1;
      o[i] = (uint32_t)u;
      sum += u;
      if (o[i] > o[best]) best = i;
    }
    if (sum > 4294967296ull) return SX_EINVAL;
    o[best] += (uint32_t)(4294967296ull - sum);
  }
  return SX_OK;
}

extern "C" int sx_mlp_proposal(sx_mlp *m, const uint16_t *ctx, uint32_t n, double eps, uint32_t *out) {
  if (!m || !ctx || n == 0 || !out) return SX_EINVAL;
  std::vector<float> probs((size_t)n * kOut), q((size_t)n * kInsts);
  int rc = sx_mlp_predict(m, ctx, n, probs.data());
  if (rc) return rc;
  for (uint32_t r = 0; r < n; r++) {
    float *qr = &q[(size_t)r * kInsts];
    const float *pr = &probs[(size_t)r * kOut];
    for (int i = 0; i < kInsts; i++) qr[i] = pr[i];
    for (int j = 0; j < kCtx; j++) {
      const uint16_t t = ctx[(size_t)r * kCtx + j];
      if (t < kInsts) qr[t] += pr[kInsts + j];  // padding positions carry no instruction
    }
  }
  return sx_proposal_weights(q.data(), n, eps, out);
}

// C = A B on the device for the MFMA layout test: A[M][K], B[K][N] bf16
// (row-major), C fp32.
extern "C" int sx_mfma_gemm_test(sx_handle *h, const uint16_t *A, const uint16_t *B, int M, int N, int K, float *C) {
  if (!h || !A || !B || !C || M <= 0 || N <= 0 || K <= 0) return SX_EINVAL;
  hipStream_t s = sx_internal_stream(h);
  uint16_t *dA = nullptr, *dB = nullptr;
  float *dC = nullptr;
  int rc = SX_OK;
  if (hipMalloc(&dA, (size_t)M * K * 2) != hipSuccess || hipMalloc(&dB, (size_t)K * N * 2) != hipSuccess ||
      hipMalloc(&dC, (size_t)M * N * 4) != hipSuccess) {
    rc = SX_ENOMEM;
  } else {
    (void)hipMemcpyAsync(dA, A, (size_t)M * K * 2, hipMemcpyHostToDevice, s);
    (void)hipMemcpyAsync(dB, B, (size_t)K * N * 2, hipMemcpyHostToDevice, s);
    hipLaunchKernelGGL(k_gemm, dim3(nb(N, 64), nb(M, 64)), dim3(256), 0, s, dA, (long)K, 1L, dB, (long)N, 1L, dC,
                       (long)N, (uint16_t *)nullptr, M, N, K, (const float *)nullptr, 0, 0);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(C, dC, (size_t)M * N * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      rc = SX_EHIP;
  }
  (void)hipFree(dA);
  (void)hipFree(dB);
  (void)hipFree(dC);
  return rc;
}
