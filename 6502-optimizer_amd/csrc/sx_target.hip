// sx_target.hip -- the `target` command on the device (sx_target_search in
// include/sx6502.h): the shortest replacement of one sequence, candidates of
// length 1..4 enumerated on the fly in Go order.
//
// Replaces search.SearchSingle(Masked) (pkg/search/search.go:76-106) and the
// per-target loops WorkerPool.processTask / processTaskMasked
// (pkg/search/worker.go:119-245): for each candidate length in turn, the first
// candidate in Go order that passes the byte filter, the pruner, QuickCheck
// and ExhaustiveCheck is the rule of the exact pass; with dead flags, the
// first that fails the exact QuickCheck but passes the masked QuickCheck, the
// masked ExhaustiveCheck and has FlagDiff != 0 is the rule of the masked pass.
//
// A candidate of length L is a prefix of L - 1 instructions and a last
// instruction.  One wave holds 8 prefixes x 8 QuickCheck vectors (lane =
// prefix * 8 + vector): each lane runs its prefix once (divergent, amortised
// over 2,863 last instructions), then the wave walks the 2,863 last
// instructions in Go order with a wave-uniform opcode (scalar dispatch), and a
// group whose 8 lanes all match is a QuickCheck survivor.  Survivors (the Go
// index of the candidate, with an exact-match bit) are compacted through an
// LDS queue per wave; the host sorts them, applies the full pruner and runs
// the complete ExhaustiveCheck in Go order until the first pass.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
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

constexpr int kWave = 64;
constexpr int kQueue = 128;       // LDS survivor queue entries per wave
constexpr int kMaxTarget = 8;
constexpr int kMaxCand = 4;
constexpr uint64_t kExactBit = 1ull << 63;

__constant__ uint8_t t_reads[OP_COUNT] = SX_READS_TABLE;
__constant__ uint8_t t_writes[OP_COUNT] = SX_WRITES_TABLE;
__constant__ uint8_t t_vec[32][8] = SX_MID_VECTORS;  // the first 8 are the QuickCheck vectors

const uint8_t h_reads[OP_COUNT] = SX_READS_TABLE;
const uint8_t h_writes[OP_COUNT] = SX_WRITES_TABLE;
const uint8_t h_vec[32][8] = SX_MID_VECTORS;

__device__ __forceinline__ uint32_t uni(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }

// Survivors of candidates prefix * 2863 + last, prefix in [p0, p0 + np).
__global__ __launch_bounds__(256) void k_target_qc(uint32_t kp, uint64_t p0, uint64_t np,
                                                   const uint32_t *__restrict__ tout, int tbytes, int model,
                                                   uint32_t pm_masked, uint64_t *__restrict__ surv,
                                                   uint32_t cap, uint32_t *__restrict__ count) {
  __shared__ uint64_t queue[4][kQueue];
  const uint32_t lane = threadIdx.x & (kWave - 1), wv = uni(threadIdx.x >> 6);  // wave-uniform
  const uint32_t g = lane >> 3, v = lane & 7;
  const uint64_t pi = p0 + ((uint64_t)blockIdx.x * 4 + wv) * 8 + g;
  bool active = pi < p0 + np;
  // decode the prefix (most significant instruction first)
  uint32_t w[kMaxCand - 1] = {0, 0, 0};
  int pbytes = 0;
  {
    uint64_t x = active ? pi : 0;
    for (int k = (int)kp - 1; k >= 0; k--) {
      w[k] = go_inst((uint32_t)(x % kNumInsts));
      x /= kNumInsts;
    }
    for (uint32_t k = 0; k < kp; k++) pbytes += byte_size(w[k] & 0xFFFFu, model);
  }
  // a prunable prefix makes every extension prunable (NOP, a dead write, a
  // stack violation or a non-canonical pair inside it stays in the candidate)
  if (kp >= 1 && should_prune(w, (int)kp, t_reads, t_writes)) active = false;
  if (pbytes + 1 >= tbytes) active = false;
  State s;
  s.a = t_vec[v][0]; s.x = t_vec[v][1]; s.y = t_vec[v][2]; s.p = t_vec[v][3];
  s.s = t_vec[v][4]; s.m = t_vec[v][5]; s.s0 = t_vec[v][6]; s.s1 = t_vec[v][7];
  for (uint32_t k = 0; k < kp; k++) step(s, w[k] & 0xFFFFu, w[k] >> 16);
  const uint32_t tlo = tout[2 * v], thi = tout[2 * v + 1];
  const uint32_t mlo = 0x00FFFFFFu | (pm_masked << 24);
  const uint32_t flo = 0x00FFFFFFu | (kPMask << 24);
  uint32_t qn = 0;

  auto flush = [&]() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(count, qn);
    base = uni(base);
    for (uint32_t i = lane; i < qn; i += kWave)
      if (base + i < cap) surv[base + i] = queue[wv][i];
    __builtin_amdgcn_wave_barrier();
    qn = 0;
  };

  if (__ballot(active) != 0) {
    for (uint32_t i = 0; i < (uint32_t)kNumInsts; i++) {
      const uint32_t wi = go_inst(i), op = wi & 0xFFFFu;
      const int bi = byte_size(op, model);
      // the last instructions are the 47 one-byte ones, then the two-byte ones
      if (i == (uint32_t)kNumNonImm && __ballot(active && pbytes + 2 < tbytes) == 0) break;
      const bool ok_b = active && pbytes + bi < tbytes;
      State s2 = s;
      step(s2, op, wi >> 16);
      const uint32_t dl = pack_lo(s2) ^ tlo, dh = pack_hi(s2) ^ thi;
      const bool okm = ok_b && ((dl & mlo) | dh) == 0;
      const uint64_t mm = __ballot(okm);
      if (mm == 0) continue;
      const uint64_t mf = __ballot(okm && ((dl & flo) | dh) == 0);
      const bool full = v == 0 && ((mm >> (8 * g)) & 0xFFu) == 0xFFu;
      const uint64_t fm = __ballot(full);
      if (fm == 0) continue;
      if (full) {
        const bool exact = ((mf >> (8 * g)) & 0xFFu) == 0xFFu;
        const uint32_t slot = qn + (uint32_t)__popcll(fm & ((1ull << lane) - 1));
        queue[wv][slot] = (pi * kNumInsts + i) | (exact ? kExactBit : 0ull);
      }
      qn += (uint32_t)__popcll(fm);
      if (qn > kQueue - 8) flush();
    }
  }
  if (qn) flush();
}

uint64_t ipow(uint64_t b, int e) {
  uint64_t r = 1;
  while (e-- > 0) r *= b;
  return r;
}

void run_seq(State &s, const uint32_t *w, int n) {
  for (int k = 0; k < n; k++) step(s, w[k] & 0xFFFFu, w[k] >> 16);
}

State vec_state(int v) {
  State s;
  s.a = h_vec[v][0]; s.x = h_vec[v][1]; s.y = h_vec[v][2]; s.p = h_vec[v][3];
  s.s = h_vec[v][4]; s.m = h_vec[v][5]; s.s0 = h_vec[v][6]; s.s1 = h_vec[v][7];
  return s;
}

// FlagDiff (pkg/search/verifier.go:591-604) on the 8 QuickCheck vectors.
uint32_t flag_diff(const uint32_t *t, int tl, const uint32_t *c, int cl) {
  uint32_t diff = 0;
  for (int v = 0; v < 8; v++) {
    State a = vec_state(v), b = vec_state(v);
    run_seq(a, t, tl);
    run_seq(b, c, cl);
    if (a.a != b.a || a.x != b.x || a.y != b.y || a.s != b.s || a.m != b.m || a.s0 != b.s0 || a.s1 != b.s1)
      return 0;
    diff |= (a.p ^ b.p) & kPMask;
  }
  return diff;
}

#define TCHK(call)                                               \
  do {                                                           \
    hipError_t e_ = (call);                                      \
    if (e_ != hipSuccess) {                                      \
      rc = sx_internal_error(h, SX_EHIP, hipGetErrorString(e_)); \
      goto done;                                                 \
    }                                                            \
  } while (0)

}  // namespace

extern "C" int sx_target_search(sx_handle *h, const uint32_t *target, int tl, int max_cand_len, uint32_t dead_flags,
                                int bytes_model, sx_target_result out[2], uint64_t *scanned) {
  if (!h || !target || !out || tl < 1 || tl > kMaxTarget || max_cand_len < 1 || max_cand_len > kMaxCand ||
      dead_flags > 0xFF || bytes_model < 0 || bytes_model > 1)
    return SX_EINVAL;
  for (int k = 0; k < tl; k++)
    if ((target[k] & 0xFFFFu) >= OP_COUNT) return sx_internal_error(h, SX_EOPCODE, "opcode out of range in target");
  memset(out, 0, 2 * sizeof(sx_target_result));
  if (scanned) *scanned = 0;
  hipStream_t stream = sx_internal_stream(h);
  int rc = SX_OK;
  int tbytes = 0, tcyc = 0;
  for (int k = 0; k < tl; k++) tbytes += byte_size(target[k] & 0xFFFFu, bytes_model), tcyc += cycles(target[k] & 0xFFFFu);
  uint32_t tout_h[16];
  for (int v = 0; v < 8; v++) {
    State s = vec_state(v);
    run_seq(s, target, tl);
    tout_h[2 * v] = pack_lo(s);
    tout_h[2 * v + 1] = pack_hi(s);
  }
  const uint32_t pm_masked = kPMask & ~dead_flags;
  uint32_t *d_tout = nullptr, *d_count = nullptr;
  uint64_t *d_surv = nullptr;
  uint32_t cap = 1u << 20;
  bool need_exact = true, need_masked = dead_flags != 0;
  std::vector<uint64_t> hs;
  TCHK(hipMalloc(&d_tout, sizeof(tout_h)));
  TCHK(hipMalloc(&d_count, 16));
  TCHK(hipMalloc(&d_surv, (size_t)cap * 8));
  TCHK(hipMemcpyAsync(d_tout, tout_h, sizeof(tout_h), hipMemcpyHostToDevice, stream));
  // each pass stops at the first candidate length with a hit (worker.go:123-177, 184-244)
  for (int cl = 1; cl <= max_cand_len && (need_exact || need_masked); cl++) {
    const uint64_t nprefix = ipow(kNumInsts, cl - 1);
    uint64_t chunk = 1u << 12;  // prefixes per launch; grows so that early hits come back quickly
    for (uint64_t p0 = 0; p0 < nprefix && (need_exact || need_masked);
         p0 += chunk, chunk = std::min<uint64_t>(chunk * 2, 1u << 21)) {
      const uint64_t np = std::min<uint64_t>(chunk, nprefix - p0);
      uint32_t n = 0;
      for (;;) {
        TCHK(hipMemsetAsync(d_count, 0, 4, stream));
        const uint64_t blocks = (np + 31) / 32;  // 32 prefixes per 256-thread block
        hipLaunchKernelGGL(k_target_qc, dim3((unsigned)blocks), dim3(256), 0, stream, (uint32_t)(cl - 1), p0, np,
                           d_tout, tbytes, bytes_model, pm_masked, d_surv, cap, d_count);
        TCHK(hipGetLastError());
        TCHK(hipMemcpyAsync(&n, d_count, 4, hipMemcpyDeviceToHost, stream));
        TCHK(hipStreamSynchronize(stream));
        if (n <= cap) break;
        cap = n + n / 4 + 1024;  // re-run with room for every survivor
        (void)hipFree(d_surv);
        d_surv = nullptr;
        TCHK(hipMalloc(&d_surv, (size_t)cap * 8));
      }
      if (scanned) *scanned += np * kNumInsts;
      if (n == 0) continue;
      hs.resize(n);
      TCHK(hipMemcpyAsync(hs.data(), d_surv, (size_t)n * 8, hipMemcpyDeviceToHost, stream));
      TCHK(hipStreamSynchronize(stream));
      std::sort(hs.begin(), hs.end(), [](uint64_t a, uint64_t b) { return (a & ~kExactBit) < (b & ~kExactBit); });
      // candidates in Go order for each open pass, after the full pruner
      for (int pass = 0; pass < 2; pass++) {
        if (pass == 0 ? !need_exact : !need_masked) continue;
        std::vector<uint32_t> tw, cw;
        std::vector<uint8_t> tlv, clv;
        for (uint64_t e : hs) {
          const bool exact = (e & kExactBit) != 0;
          if (pass == 0 ? !exact : exact) continue;  // the masked pass takes candidates failing the exact check
          uint64_t x = e & ~kExactBit;
          uint32_t c[kMaxCand];
          for (int k = cl - 1; k >= 0; k--) {
            c[k] = go_inst((uint32_t)(x % kNumInsts));
            x /= kNumInsts;
          }
          if (should_prune(c, cl, h_reads, h_writes)) continue;
          if (pass == 1 && flag_diff(target, tl, c, cl) == 0) continue;
          for (int k = 0; k < 10; k++) tw.push_back(k < tl ? target[k] : 0u);
          for (int k = 0; k < 10; k++) cw.push_back(k < cl ? c[k] : 0u);
          tlv.push_back((uint8_t)tl);
          clv.push_back((uint8_t)cl);
        }
        const uint64_t m = clv.size();
        // verify in Go order, in growing batches, until the first pass
        for (uint64_t b0 = 0, bs = 64; b0 < m; b0 += bs, bs = std::min<uint64_t>(bs * 4, 1u << 16)) {
          const uint64_t nb = std::min<uint64_t>(bs, m - b0);
          std::vector<uint8_t> ok(nb);
          rc = sx_verify_seqs(h, tw.data() + b0 * 10, tlv.data() + b0, cw.data() + b0 * 10, clv.data() + b0, nb,
                              pass == 0 ? 0u : dead_flags, ok.data());
          if (rc) goto done;
          uint64_t hit = nb;
          for (uint64_t i = 0; i < nb; i++)
            if (ok[i]) {
              hit = i;
              break;
            }
          if (hit == nb) continue;
          sx_target_result &r = out[pass];
          const uint32_t *c = cw.data() + (b0 + hit) * 10;
          int cb = 0, cc = 0;
          for (int k = 0; k < cl; k++) {
            r.rep[k] = c[k];
            cb += byte_size(c[k] & 0xFFFFu, bytes_model);
            cc += cycles(c[k] & 0xFFFFu);
          }
          r.rep_len = (uint8_t)cl;
          r.found = 1;
          r.pass = (uint8_t)(pass + 1);
          r.dead_flags = (uint8_t)(pass == 1 ? flag_diff(target, tl, c, cl) : 0);
          r.bytes_saved = (int16_t)(tbytes - cb);
          r.cycles_saved = (int16_t)(tcyc - cc);
          if (pass == 0) need_exact = false; else need_masked = false;
          break;
        }
      }
    }
  }
done:
  if (d_tout) (void)hipFree(d_tout);
  if (d_count) (void)hipFree(d_count);
  if (d_surv) (void)hipFree(d_surv);
  return rc;
}
