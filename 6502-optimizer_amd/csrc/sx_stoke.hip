// sx_stoke.hip -- STOKE search on the device (sx_stoke_run in include/sx6502.h).
//
// One wave per Markov chain.  Lane v < 32 holds Mid vector v and the
// target's output on it; a proposal is evaluated by running the candidate
// (wave-uniform instructions, read from LDS) on every lane's vector and
// reducing the Hamming distance across lanes.  Proposal and acceptance are
// wave-uniform scalar work: Philox4x32-10 per chain (counter = step, chain,
// stream), the move set of pkg/stoke/mutator.go:30-106 with weights
// 102/52/51/25/26 out of 256, instruction draws from integer weights that sum
// to 2^32 (uniform or learned), and Metropolis-Hastings acceptance with the
// exact proposal ratio.  A chain whose accepted state matches the target on
// all 32 vectors with fewer bytes posts it for verification (up to kPosts
// distinct candidates per chain and launch, so one candidate that matches on
// the vectors but fails the sweep does not hide the next); the host checks
// posted candidates with the complete ExhaustiveCheck between launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <set>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/sx6502.h"
#include "sx_isa.h"

using namespace sx;

hipStream_t sx_internal_stream(sx_handle *h);
int sx_internal_error(sx_handle *h, int code, const char *msg);
int sx_verify_seqs(sx_handle *h, const uint32_t *t_words, const uint8_t *t_len, const uint32_t *c_words,
                   const uint8_t *c_len, uint64_t n, uint32_t dead_flags, uint8_t *ok);

namespace {

constexpr int kMaxSeq = 10;
constexpr int kWave = 64;
constexpr int kPosts = 4;  // post slots per chain

__constant__ uint8_t s_vec[32][8] = SX_MID_VECTORS;

struct ChainState {
  uint32_t seq[kMaxSeq];
  int32_t len;
  int32_t cost10;
  double temp;
  uint64_t step;
  uint64_t accepted;
  uint32_t posted;   // filled post slots (0..kPosts)
  uint32_t pad;
};

struct Post {
  uint32_t seq[kMaxSeq];
  int32_t len;
  uint32_t pad;
  uint64_t step;
};

__device__ __forceinline__ uint32_t uni(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }

__device__ __forceinline__ void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                       uint32_t k1, uint32_t out[4]) {
#pragma unroll
  for (int r = 0; r < 10; r++) {
    uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
    uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
    uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

// Go-order index of an instruction word (pkg/search/enumerator.go:14-38).
__device__ __forceinline__ uint32_t go_index(uint32_t w) {
  uint32_t op = w & 0xFFFFu;
  if (op < 26) return op;
  if (op >= LDA_M) return op - LDA_M + 26;
  return kNumNonImm + (op - LDA_IMM) * 256 + ((w >> 16) & 0xFFu);
}

__device__ __forceinline__ uint32_t go_inst_d(uint32_t i) {
  if (i < 26) return i;
  if (i < (uint32_t)kNumNonImm) return i - 26 + LDA_M;
  uint32_t k = i - kNumNonImm;
  return (LDA_IMM + (k >> 8)) | ((k & 0xFFu) << 16);
}

__device__ __forceinline__ uint32_t mulhi32(uint32_t a, uint32_t n) { return __umulhi(a, n); }

// Instruction tables.  The uniform table is 2^32 / 2863 per instruction, one
// more for the first 2^32 mod 2863, so both its draw and its weights are
// closed forms and need no memory.  A per-target table is searched in two
// wave-wide steps instead of a 12-step dependent binary search: lane k keeps
// the last cumulative weight of segment k (64 segments of 45) in a register
// for the whole launch, so a ballot finds the segment, and one coalesced load
// of that segment and a second ballot find the entry.
constexpr uint32_t kSeg = 45;  // 64 x 45 = 2,880 >= 2,863
constexpr uint32_t kUniBase = (uint32_t)(4294967296ull / kNumInsts);
constexpr uint32_t kUniExtra = (uint32_t)(4294967296ull % kNumInsts);

struct Table {
  const uint64_t *cum;  // inclusive cumulative weights (cum[2862] = 2^32), null = uniform
  const uint32_t *w;
  uint64_t pivot;       // this lane's segment end

  // the smallest i with r < cum[i]; called with all lanes active
  __device__ __forceinline__ uint32_t draw(uint32_t r, uint32_t lane) const {
    if (!cum) {
      const uint64_t hi = (uint64_t)kUniExtra * (kUniBase + 1);  // draws below hi land on a heavier entry
      return (uint64_t)r < hi ? r / (kUniBase + 1) : kUniExtra + (uint32_t)(((uint64_t)r - hi) / kUniBase);
    }
    const uint64_t b1 = __ballot((uint64_t)r < pivot);
    const uint32_t k0 = (uint32_t)__builtin_ctzll(b1);  // nonzero: the last entry is 2^32 > r
    const uint32_t idx = k0 * kSeg + lane;
    const bool in_seg = lane < kSeg && idx < (uint32_t)kNumInsts;
    const uint64_t b2 = __ballot(in_seg && (uint64_t)r < cum[in_seg ? idx : 0]);
    return uni(k0 * kSeg + (uint32_t)__builtin_ctzll(b2));
  }
  // q(i) = weight / 2^32
  __device__ __forceinline__ double q(uint32_t i) const {
    const uint32_t wi = cum ? w[i] : kUniBase + (i < kUniExtra ? 1u : 0u);
    return wi * (1.0 / 4294967296.0);
  }
};

constexpr double W_REP = 102.0 / 256.0, W_DEL = 51.0 / 256.0, W_INS = 25.0 / 256.0, W_IMM = 26.0 / 256.0;

// chains: n; targets: words [nt][8], lens, per-chain target index = c / cpt.
// cum: per weight table, 2863 uint64 cumulative weights (table index =
// target index when per_target, else 0).
__global__ __launch_bounds__(256) void k_stoke(ChainState *chains, uint32_t n, uint32_t cpt,
                                               const uint32_t *__restrict__ twords,
                                               const uint8_t *__restrict__ tlens,
                                               const uint64_t *__restrict__ cum,
                                               const uint32_t *__restrict__ weights, int per_target,
                                               uint32_t steps, uint32_t k0, uint32_t k1, double decay,
                                               uint32_t pmask, const uint32_t *__restrict__ stop,
                                               Post *posts) {
  __shared__ uint32_t xs[4][kMaxSeq + 1];
  __shared__ uint32_t ys[4][kMaxSeq + 1];
  __shared__ uint32_t ps[4][kMaxSeq + 1];  // last posted candidate (words, then length)
  const uint32_t lane = threadIdx.x & (kWave - 1), wv = uni(threadIdx.x >> 6);  // wave-uniform
  const uint32_t c = blockIdx.x * 4 + wv;
  if (c >= n) return;
  const uint32_t ti = uni(c / cpt);
  ChainState *cs = &chains[c];
  const uint32_t tl = uni(tlens[ti]);
  uint32_t tw[8];
  int tbytes = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    tw[k] = uni(twords[ti * 8 + k]);
    if (k < (int)tl) tbytes += byte_size(tw[k] & 0xFFFFu, 0);
  }
  const int max_len = (int)tl + 2 > kMaxSeq ? kMaxSeq : ((int)tl + 2 > 10 ? (int)tl + 2 : 10);
  // target output on this lane's vector
  State base;
  const uint32_t v = lane & 31u;
  base.a = s_vec[v][0]; base.x = s_vec[v][1]; base.y = s_vec[v][2]; base.p = s_vec[v][3];
  base.s = s_vec[v][4]; base.m = s_vec[v][5]; base.s0 = s_vec[v][6]; base.s1 = s_vec[v][7];
  State to = base;
  for (uint32_t k = 0; k < tl; k++) step(to, tw[k] & 0xFFFFu, tw[k] >> 16);
  const uint32_t tlo = pack_lo(to), thi = pack_hi(to);
  const uint32_t plo = 0x00FFFFFFu | (pmask << 24);
  Table tab;
  tab.cum = per_target ? cum + (uint64_t)ti * kNumInsts : nullptr;
  tab.w = per_target ? weights + (uint64_t)ti * kNumInsts : nullptr;
  tab.pivot = per_target ? tab.cum[min(lane * kSeg + kSeg - 1, (uint32_t)kNumInsts - 1)] : 0;

  if (lane < kMaxSeq) xs[wv][lane] = cs->seq[lane];
  int L = (int)uni((uint32_t)cs->len);
  int cost10 = (int)uni((uint32_t)cs->cost10);
  double temp = cs->temp;
  uint64_t st = cs->step;
  uint64_t acc = cs->accepted;
  uint32_t posted = uni(cs->posted);
  if (lane <= kMaxSeq) ps[wv][lane] = 0xFFFFFFFFu;  // no post yet in this launch
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();

  for (uint32_t it = 0; it < steps; it++) {
    if ((it & 63u) == 0 && uni(((volatile const uint32_t *)stop)[ti])) break;
    const uint32_t s_lo = uni((uint32_t)st), s_hi = uni((uint32_t)(st >> 32));
    uint32_t r[4], a[4];
    philox(s_lo, s_hi, c, 0, k0, k1, r);
    philox(s_lo, s_hi, c, 1, k0, k1, a);
    for (int k = 0; k < 4; k++) {
      r[k] = uni(r[k]);
      a[k] = uni(a[k]);
    }
    const uint32_t m = r[0] >> 24;
    // current sequence words (uniform)
    uint32_t x[kMaxSeq];
#pragma unroll
    for (int k = 0; k < kMaxSeq; k++) x[k] = uni(xs[wv][k]);
    int nimm = 0;
#pragma unroll
    for (int k = 0; k < kMaxSeq; k++) nimm += (k < L && has_imm(x[k] & 0xFFFFu)) ? 1 : 0;
    // proposal: kind 0 none, 1 replace at pos by v, 2 swap pos/pos+1, 3 delete pos, 4 insert v at pos
    int kind = 0, pos = 0, nl = L;
    uint32_t vnew = 0;
    double ratio = 1.0;
    if (m < 102u) {
      pos = (int)mulhi32(r[1], (uint32_t)L);
      vnew = go_inst_d(tab.draw(r[2], lane));
      uint32_t old = x[0];
#pragma unroll
      for (int k = 1; k < kMaxSeq; k++) if (k == pos) old = x[k];
      if (vnew != old) {
        double same = ((vnew & 0xFFFFu) == (old & 0xFFFFu) && has_imm(old & 0xFFFFu)) ? W_IMM / (nimm * 256.0) : 0.0;
        double qo = tab.q(go_index(old)), qn = tab.q(go_index(vnew));
        ratio = (W_REP * qo / L + same) / (W_REP * qn / L + same);
        kind = 1;
      }
    } else if (m < 154u) {
      if (L >= 2) {
        pos = (int)mulhi32(r[1], (uint32_t)(L - 1));
        uint32_t p0 = x[0], p1 = x[1];
#pragma unroll
        for (int k = 0; k + 1 < kMaxSeq; k++) if (k == pos) { p0 = x[k]; p1 = x[k + 1]; }
        if (p0 != p1) kind = 2;
      }
    } else if (m < 205u) {
      if (L > 1) {
        pos = (int)mulhi32(r[1], (uint32_t)L);
        uint32_t old = x[0];
#pragma unroll
        for (int k = 1; k < kMaxSeq; k++) if (k == pos) old = x[k];
        ratio = W_INS * tab.q(go_index(old)) / W_DEL;
        kind = 3;
        nl = L - 1;
      }
    } else if (m < 230u) {
      if (L < max_len) {
        pos = (int)mulhi32(r[1], (uint32_t)(L + 1));
        vnew = go_inst_d(tab.draw(r[2], lane));
        ratio = W_DEL / (W_INS * tab.q(go_index(vnew)));
        kind = 4;
        nl = L + 1;
      }
    } else {
      if (nimm > 0) {
        int kth = (int)mulhi32(r[1], (uint32_t)nimm);
        int p = 0, seen = 0;
#pragma unroll
        for (int k = 0; k < kMaxSeq; k++) {
          bool im = k < L && has_imm(x[k] & 0xFFFFu);
          if (im && seen == kth) p = k;
          seen += im ? 1 : 0;
        }
        pos = p;
        uint32_t old = x[0];
#pragma unroll
        for (int k = 1; k < kMaxSeq; k++) if (k == pos) old = x[k];
        vnew = (old & 0xFFFFu) | ((r[3] >> 24) << 16);
        if (vnew != old) {
          double same = W_IMM / (nimm * 256.0);
          double qo = tab.q(go_index(old)), qn = tab.q(go_index(vnew));
          ratio = (W_REP * qo / L + same) / (W_REP * qn / L + same);
          kind = 1;
        }
      }
    }
    if (kind != 0) {
      // build the proposal in LDS, lane k writes position k
      if (lane < kMaxSeq) {
        const int k = (int)lane;
        uint32_t y = xs[wv][k];
        if (kind == 1 && k == pos) y = vnew;
        if (kind == 2 && k == pos) y = xs[wv][k + 1];
        if (kind == 2 && k == pos + 1) y = xs[wv][k - 1];
        if (kind == 3 && k >= pos) y = k + 1 < kMaxSeq ? xs[wv][k + 1] : 0u;
        if (kind == 4 && k > pos) y = xs[wv][k - 1];
        if (kind == 4 && k == pos) y = vnew;
        ys[wv][k] = y;
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      // evaluate on the 32 vectors
      State sv = base;
      int bytes = 0;
      for (int k = 0; k < nl; k++) {
        uint32_t w = uni(ys[wv][k]);
        step(sv, w & 0xFFFFu, w >> 16);
        bytes += byte_size(w & 0xFFFFu, 0);
      }
      uint32_t h = lane < 32 ? __popc((pack_lo(sv) ^ tlo) & plo) + __popc(pack_hi(sv) ^ thi) : 0u;
      for (int off = 32; off >= 1; off >>= 1) h += __shfl_xor(h, off);
      h = uni(h);
      const int c10 = 10 * (int)h + bytes;
      const double la = -((double)(c10 - cost10) / 10.0) / temp + log(ratio);
      const double u = ((double)(a[0] >> 5) * 67108864.0 + (double)(a[1] >> 6) + 0.5) / 9007199254740992.0;
      const bool accept = la >= 0.0 || u < exp(la);
      if (uni(accept ? 1u : 0u)) {
        if (lane < kMaxSeq) xs[wv][lane] = ys[wv][lane];
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        L = nl;
        cost10 = c10;
        acc++;
        if (h == 0 && bytes < tbytes && posted < (uint32_t)kPosts) {
          // skip a repeat of the last posted candidate
          const bool same_lane = lane < kMaxSeq ? (lane >= (uint32_t)nl || ps[wv][lane] == ys[wv][lane]) : true;
          const bool repeat = __all(same_lane) && ps[wv][kMaxSeq] == (uint32_t)nl;
          if (!repeat) {
            Post *pp = &posts[(uint64_t)c * kPosts + posted];
            if (lane < kMaxSeq) {
              pp->seq[lane] = ys[wv][lane];
              ps[wv][lane] = ys[wv][lane];
            }
            if (lane == 0) {
              pp->len = nl;
              pp->step = st;
              ps[wv][kMaxSeq] = (uint32_t)nl;
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            posted++;
          }
        }
      }
    }
    temp *= decay;
    st++;
  }
  if (lane < kMaxSeq) cs->seq[lane] = xs[wv][lane];
  if (lane == 0) {
    cs->len = L;
    cs->cost10 = cost10;
    cs->temp = temp;
    cs->step = st;
    cs->accepted = acc;
    cs->posted = posted;
  }
}

// Initial chain states: every chain starts at its target.
__global__ void k_stoke_init(ChainState *chains, uint32_t n, uint32_t cpt, const uint32_t *twords,
                             const uint8_t *tlens, double t0) {
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  uint32_t ti = c / cpt;
  ChainState s;
  int bytes = 0;
  for (int k = 0; k < kMaxSeq; k++) s.seq[k] = k < tlens[ti] ? twords[ti * 8 + k] : 0u;
  for (int k = 0; k < tlens[ti]; k++) bytes += byte_size(twords[ti * 8 + k] & 0xFFFFu, 0);
  s.len = tlens[ti];
  s.cost10 = bytes;  // Hamming 0 at the target
  s.temp = t0;
  s.step = 0;
  s.accepted = 0;
  s.posted = 0;
  s.pad = 0;
  chains[c] = s;
}

void uniform_weights(std::vector<uint32_t> &w) {
  const uint64_t total = 1ull << 32;
  const uint32_t base = (uint32_t)(total / kNumInsts), extra = (uint32_t)(total % kNumInsts);
  w.resize(kNumInsts);
  for (int i = 0; i < kNumInsts; i++) w[i] = base + (i < (int)extra ? 1u : 0u);
}

#define HCHK(call)                                                   \
  do {                                                               \
    hipError_t e_ = (call);                                          \
    if (e_ != hipSuccess) {                                          \
      rc = sx_internal_error(h, SX_EHIP, hipGetErrorString(e_));     \
      goto done;                                                     \
    }                                                                \
  } while (0)

// FlagDiff (pkg/search/verifier.go:591-604) on the host: the P bits in which
// the two sequences' outputs differ over the 8 QuickCheck vectors, 0 when
// another byte differs.
uint32_t host_flag_diff(const uint32_t *t, int tl, const uint32_t *c, int cl) {
  static const uint8_t vec[32][8] = SX_MID_VECTORS;
  uint32_t d = 0;
  for (int v = 0; v < 8; v++) {
    State a = unpack(vec[v][0] | vec[v][1] << 8 | vec[v][2] << 16 | (uint32_t)vec[v][3] << 24,
                     vec[v][4] | vec[v][5] << 8 | vec[v][6] << 16 | (uint32_t)vec[v][7] << 24);
    State b = a;
    for (int k = 0; k < tl; k++) step(a, t[k] & 0xFFFFu, t[k] >> 16);
    for (int k = 0; k < cl; k++) step(b, c[k] & 0xFFFFu, c[k] >> 16);
    if (((pack_lo(a) ^ pack_lo(b)) & 0x00FFFFFFu) || pack_hi(a) != pack_hi(b)) return 0;
    d |= ((a.p ^ b.p) & kPMask);
  }
  return d;
}

// The chain state a chain starts (and, collecting, restarts) from: its
// target at temperature t0 (pkg/stoke/mcmc.go:28-47); the step counter runs on.
ChainState start_state(const uint32_t *targets, const uint8_t *lens, uint32_t ti, double t0, uint64_t step,
                       uint64_t accepted) {
  ChainState s;
  memset(&s, 0, sizeof(s));
  int bytes = 0;
  for (int k = 0; k < kMaxSeq; k++) s.seq[k] = k < lens[ti] ? targets[ti * 8 + k] : 0u;
  for (int k = 0; k < lens[ti]; k++) bytes += byte_size(targets[ti * 8 + k] & 0xFFFFu, 0);
  s.len = lens[ti];
  s.cost10 = bytes;
  s.temp = t0;
  s.step = step;
  s.accepted = accepted;
  return s;
}

// Both STOKE entry points.  first: out[n] gets each target's first verified
// rewrite (smallest step, then chain, among the launch's verified posts).
// found != null (collect, pkg/stoke/search.go:101-153 with Deduplicate,
// :179-190): every verified rewrite is kept once per (target, replacement) in
// the order found (launch, then target, step, chain), and the chain that
// posted it restarts from its target at t0 at the next launch.
int stoke_core(sx_handle *h, const sx_stoke_cfg *cfg, const uint32_t *targets, const uint8_t *lens, uint64_t n,
               sx_stoke_result *out, std::vector<sx_stoke_found> *found, sx_stoke_chain *chains_out) {
  if (!h || !cfg || (n && (!targets || !lens)) || (n && !out && !found)) return SX_EINVAL;
  if (cfg->struct_size != sizeof(sx_stoke_cfg)) return sx_internal_error(h, SX_EINVAL, "sx_stoke_cfg size mismatch");
  if (cfg->chains_per_target == 0 || cfg->dead_flags > 0xFF || !(cfg->t0 > 0) || !(cfg->decay > 0) ||
      cfg->decay > 1.0)
    return sx_internal_error(h, SX_EINVAL, "bad STOKE configuration");
  for (uint64_t i = 0; i < n; i++) {
    if (lens[i] < 1 || lens[i] > 8) return sx_internal_error(h, SX_EINVAL, "target length must be 1..8");
    for (int k = 0; k < lens[i]; k++)
      if ((targets[i * 8 + k] & 0xFFFFu) >= OP_COUNT) return sx_internal_error(h, SX_EOPCODE, "opcode out of range");
  }
  if (out) memset(out, 0, n * sizeof(sx_stoke_result));
  if (n == 0) return SX_OK;
  const bool collect = found != nullptr;
  const uint64_t nch64 = n * (uint64_t)cfg->chains_per_target;
  if (nch64 > (1u << 26)) return sx_internal_error(h, SX_EINVAL, "too many chains");
  const uint32_t nch = (uint32_t)nch64, cpt = cfg->chains_per_target;
  hipStream_t stream = sx_internal_stream(h);
  int rc = SX_OK;
  ChainState *d_ch = nullptr;
  Post *d_post = nullptr;
  uint32_t *d_tw = nullptr, *d_w = nullptr, *d_stop = nullptr;
  uint8_t *d_tl = nullptr;
  uint64_t *d_cum = nullptr;
  const int per_target = cfg->weights != nullptr;
  const uint64_t ntab = per_target ? n : 1;
  std::vector<uint32_t> w;
  std::vector<uint64_t> cumh;
  std::vector<ChainState> hch(nch);
  std::vector<Post> hpost((size_t)nch * kPosts);
  std::vector<uint32_t> stoph(n, 0);
  // verdicts of (target, candidate) pairs already verified: a repeat is not swept again
  std::map<std::vector<uint32_t>, bool> verdict;
  std::set<std::vector<uint32_t>> kept;  // collect: (target, replacement) pairs already reported
  auto t_start = std::chrono::steady_clock::now();
  if (per_target) {
    w.assign(cfg->weights, cfg->weights + ntab * kNumInsts);
    for (uint64_t t = 0; t < ntab; t++) {
      uint64_t s = 0;
      for (int i = 0; i < kNumInsts; i++) s += w[t * kNumInsts + i];
      if (s != (1ull << 32)) {
        rc = sx_internal_error(h, SX_EINVAL, "proposal weights must sum to 2^32 per target");
        goto done;
      }
    }
  } else {
    uniform_weights(w);
  }
  cumh.resize(ntab * kNumInsts);
  for (uint64_t t = 0; t < ntab; t++) {
    uint64_t s = 0;
    for (int i = 0; i < kNumInsts; i++) cumh[t * kNumInsts + i] = (s += w[t * kNumInsts + i]);
  }
  HCHK(hipMalloc(&d_ch, nch * sizeof(ChainState)));
  HCHK(hipMalloc(&d_post, (size_t)nch * kPosts * sizeof(Post)));
  HCHK(hipMalloc(&d_tw, n * 8 * sizeof(uint32_t)));
  HCHK(hipMalloc(&d_tl, n));
  HCHK(hipMalloc(&d_w, ntab * kNumInsts * sizeof(uint32_t)));
  HCHK(hipMalloc(&d_cum, ntab * kNumInsts * sizeof(uint64_t)));
  HCHK(hipMalloc(&d_stop, n * sizeof(uint32_t)));
  HCHK(hipMemcpyAsync(d_tw, targets, n * 8 * 4, hipMemcpyHostToDevice, stream));
  HCHK(hipMemcpyAsync(d_tl, lens, n, hipMemcpyHostToDevice, stream));
  HCHK(hipMemcpyAsync(d_w, w.data(), w.size() * 4, hipMemcpyHostToDevice, stream));
  HCHK(hipMemcpyAsync(d_cum, cumh.data(), cumh.size() * 8, hipMemcpyHostToDevice, stream));
  HCHK(hipMemsetAsync(d_stop, 0, n * 4, stream));
  hipLaunchKernelGGL(k_stoke_init, dim3((nch + 255) / 256), dim3(256), 0, stream, d_ch, nch, cpt, d_tw, d_tl,
                     cfg->t0);
  HCHK(hipGetLastError());
  HCHK(hipStreamSynchronize(stream));
  t_start = std::chrono::steady_clock::now();
  {
    // launch lengths: the caller's, else 256 steps doubling up to 4,096 (posts
    // are verified between launches, so a short first launch reports an early
    // find early, and long later launches keep the launch overhead small)
    uint32_t per_launch = cfg->steps_per_launch ? cfg->steps_per_launch : 256;
    const uint32_t pmask = kPMask & ~cfg->dead_flags;
    uint64_t remaining = cfg->steps;
    uint32_t open = (uint32_t)n;
    while (remaining > 0 && open > 0) {
      const uint32_t s = (uint32_t)std::min<uint64_t>(remaining, per_launch);
      hipLaunchKernelGGL(k_stoke, dim3((nch + 3) / 4), dim3(256), 0, stream, d_ch, nch, cpt, d_tw, d_tl, d_cum, d_w,
                         per_target, s, (uint32_t)cfg->seed, (uint32_t)(cfg->seed >> 32), cfg->decay, pmask, d_stop,
                         d_post);
      HCHK(hipGetLastError());
      HCHK(hipMemcpyAsync(hch.data(), d_ch, nch * sizeof(ChainState), hipMemcpyDeviceToHost, stream));
      HCHK(hipStreamSynchronize(stream));
      remaining -= s;
      if (!cfg->steps_per_launch) per_launch = std::min<uint32_t>(per_launch * 2, 4096);
      std::vector<uint32_t> cand;
      for (uint32_t c = 0; c < nch; c++)
        if (hch[c].posted) cand.push_back(c);
      if (cand.empty()) continue;
      HCHK(hipMemcpyAsync(hpost.data(), d_post, (size_t)nch * kPosts * sizeof(Post), hipMemcpyDeviceToHost, stream));
      HCHK(hipStreamSynchronize(stream));
      // the launch's posts in (target, step, chain) order; pairs not yet verified go to one batch
      std::vector<uint64_t> ids;  // post slot index c * kPosts + k
      for (uint32_t c : cand) {
        const uint32_t ti = c / cpt;
        if (!collect && (stoph[ti] || out[ti].verified)) continue;  // only a target's first verified rewrite
        for (uint32_t k = 0; k < hch[c].posted && k < (uint32_t)kPosts; k++) ids.push_back((uint64_t)c * kPosts + k);
      }
      std::sort(ids.begin(), ids.end(), [&](uint64_t a, uint64_t b) {
        const uint32_t ta = (uint32_t)(a / kPosts) / cpt, tb = (uint32_t)(b / kPosts) / cpt;
        if (ta != tb) return ta < tb;
        if (hpost[a].step != hpost[b].step) return hpost[a].step < hpost[b].step;
        return a < b;
      });
      auto key_of = [&](uint64_t id) {
        const Post &pp = hpost[id];
        std::vector<uint32_t> key(pp.seq, pp.seq + pp.len);
        key.push_back(0xFFFFFFFFu);  // separator: a key is (replacement, target)
        key.push_back((uint32_t)(id / kPosts) / cpt);
        return key;
      };
      {
        std::vector<uint32_t> vt, vc;
        std::vector<uint8_t> vtl, vcl;
        std::vector<std::vector<uint32_t>> vkeys;
        std::set<std::vector<uint32_t>> queued;
        for (uint64_t id : ids) {
          auto key = key_of(id);
          if (verdict.count(key) || !queued.insert(key).second) continue;
          const Post &pp = hpost[id];
          const uint32_t ti = (uint32_t)(id / kPosts) / cpt;
          for (int q = 0; q < kMaxSeq; q++) vt.push_back(q < lens[ti] ? targets[ti * 8 + q] : 0u);
          for (int q = 0; q < kMaxSeq; q++) vc.push_back(q < pp.len ? pp.seq[q] : 0u);
          vtl.push_back(lens[ti]);
          vcl.push_back((uint8_t)pp.len);
          vkeys.push_back(key);
        }
        if (!vkeys.empty()) {
          std::vector<uint8_t> res(vkeys.size());
          rc = sx_verify_seqs(h, vt.data(), vtl.data(), vc.data(), vcl.data(), vkeys.size(), cfg->dead_flags,
                              res.data());
          if (rc) goto done;
          for (size_t i = 0; i < vkeys.size(); i++) verdict[vkeys[i]] = res[i] != 0;
        }
      }
      const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
      std::vector<uint8_t> restart(collect ? nch : 0, 0);
      for (uint64_t id : ids) {
        if (!verdict[key_of(id)]) continue;
        const Post &pp = hpost[id];
        const uint32_t c = (uint32_t)(id / kPosts), ti = c / cpt;
        int tb = 0, tc = 0, cb = 0, cc = 0;
        for (int k = 0; k < lens[ti]; k++)
          tb += byte_size(targets[ti * 8 + k] & 0xFFFFu, 0), tc += cycles(targets[ti * 8 + k] & 0xFFFFu);
        for (int k = 0; k < pp.len; k++) cb += byte_size(pp.seq[k] & 0xFFFFu, 0), cc += cycles(pp.seq[k] & 0xFFFFu);
        // search.go:106-110: with dead flags the rule's mask is FlagDiff(target, best)
        const uint32_t fd = cfg->dead_flags ? host_flag_diff(targets + ti * 8, lens[ti], pp.seq, pp.len) : 0u;
        if (collect) {
          restart[c] = 1;
          if (!kept.insert(key_of(id)).second) continue;
          sx_stoke_found f;
          memset(&f, 0, sizeof(f));
          f.target = ti;
          f.chain = c;
          f.step = pp.step;
          f.seconds = secs;
          for (int k = 0; k < kMaxSeq; k++) f.rep[k] = k < pp.len ? pp.seq[k] : 0u;
          f.rep_len = (uint8_t)pp.len;
          f.dead_flags = (uint8_t)fd;
          f.bytes_saved = (int16_t)(tb - cb);
          f.cycles_saved = (int16_t)(tc - cc);
          found->push_back(f);
          continue;
        }
        sx_stoke_result &o = out[ti];
        if (o.verified) continue;  // ids are in (target, step, chain) order: the first verified post wins
        o.verified = 1;
        o.rep_len = (uint8_t)pp.len;
        for (int k = 0; k < kMaxSeq; k++) o.rep[k] = k < pp.len ? pp.seq[k] : 0u;
        o.dead_flags = (uint8_t)fd;
        o.bytes_saved = (int16_t)(tb - cb);
        o.cycles_saved = (int16_t)(tc - cc);
        o.chain = c;
        o.step = pp.step;
        o.seconds = secs;
        if (cfg->stop_when_verified) {
          stoph[ti] = 1;
          open--;
        }
      }
      // clear every posted flag so chains can post again; restart the chains that found a rewrite
      for (uint32_t c : cand) {
        hch[c].posted = 0;
        if (collect && restart[c]) hch[c] = start_state(targets, lens, c / cpt, cfg->t0, hch[c].step, hch[c].accepted);
      }
      HCHK(hipMemcpyAsync(d_ch, hch.data(), nch * sizeof(ChainState), hipMemcpyHostToDevice, stream));
      if (!collect && cfg->stop_when_verified)
        HCHK(hipMemcpyAsync(d_stop, stoph.data(), n * 4, hipMemcpyHostToDevice, stream));
      HCHK(hipStreamSynchronize(stream));
    }
  }
  if (chains_out) {
    for (uint32_t c = 0; c < nch; c++) {
      sx_stoke_chain &o = chains_out[c];
      for (int k = 0; k < kMaxSeq; k++) o.seq[k] = hch[c].seq[k];
      o.len = hch[c].len;
      o.cost10 = hch[c].cost10;
      o.step = hch[c].step;
      o.accepted = hch[c].accepted;
    }
  }
done:
  if (d_ch) (void)hipFree(d_ch);
  if (d_post) (void)hipFree(d_post);
  if (d_tw) (void)hipFree(d_tw);
  if (d_tl) (void)hipFree(d_tl);
  if (d_w) (void)hipFree(d_w);
  if (d_cum) (void)hipFree(d_cum);
  if (d_stop) (void)hipFree(d_stop);
  return rc;
}

}  // namespace

extern "C" int sx_stoke_run(sx_handle *h, const sx_stoke_cfg *cfg, const uint32_t *targets, const uint8_t *lens,
                            uint64_t n, sx_stoke_result *out, sx_stoke_chain *chains_out) {
  if (n && !out) return SX_EINVAL;
  return stoke_core(h, cfg, targets, lens, n, out, nullptr, chains_out);
}

extern "C" int sx_stoke_collect(sx_handle *h, const sx_stoke_cfg *cfg, const uint32_t *targets, const uint8_t *lens,
                                uint64_t n, sx_stoke_found *found, uint64_t cap, uint64_t *n_found,
                                sx_stoke_chain *chains_out) {
  if (!n_found || (cap && !found)) return SX_EINVAL;
  *n_found = 0;
  std::vector<sx_stoke_found> all;
  int rc = stoke_core(h, cfg, targets, lens, n, nullptr, &all, chains_out);
  if (rc) return rc;
  *n_found = all.size();
  if (cap) memcpy(found, all.data(), std::min<uint64_t>(cap, all.size()) * sizeof(sx_stoke_found));
  return SX_OK;
}
