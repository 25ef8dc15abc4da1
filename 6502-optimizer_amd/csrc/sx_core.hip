// sx_core.hip -- lib6502core: the MI355X search core of the 6502
// superoptimizer, behind the C ABI of include/sx6502.h.
//
// Pipeline of one enumeration (per segment of up to 2^22 targets):
//   1. fingerprints  8 sequences per wave in turn, lane v = Mid vector v
//                    (0..31): the sequence is wave-uniform, so the
//                    interpreter runs with scalar dispatch.  Targets and
//                    candidates get the outputs on all 32 vectors
//                    (QuickCheck vectors = 0..7); targets also get a 32-bit
//                    filter hash of their masked outputs on vectors 0..3.
//   2. QuickCheck    a lane per candidate holds its 8 QC outputs and its hash
//                    in VGPRs and walks a tile of wave-uniform targets
//                    (scalar loads of 8 hashes at a time); a hash match gets
//                    the exact 8-vector compare; survivors are collected by
//                    ballot + mbcnt into an LDS queue per wave and flushed
//                    with one atomic per batch.
//   3. MidCheck      a lane per survivor: byte filter, candidate pruner and a
//                    compare of the 24 further vector outputs; output
//                    reserved once per 1,024-thread block.
//   4. Exhaustive    jobs = (target, candidate) pairs, planned where they are
//                    made (sub-sweeps, 16K-state chunk counts); a scan turns
//                    counts into chunk offsets and a chunk -> job map, and a
//                    wave per chunk sweeps it with the packed interpreter
//                    (two states per register), early exit per job.
//                    First-hit semantics run one round (the lowest-ranked
//                    untested candidate per target and pass) and then the
//                    rest as one batch whose minimal passing rank per target
//                    and pass is the rule.
//   5. Rules         compacted in (target, pass) order on the device and
//                    copied to a pinned host buffer per segment on a second
//                    stream.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdarg>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/sx6502.h"
#include "sx_isa.h"
#include "sx_plan.h"
#include "sx_pk.h"

using namespace sx;

// ---------------------------------------------------------------- tables --

__constant__ uint8_t c_reads[OP_COUNT] = SX_READS_TABLE;
__constant__ uint8_t c_writes[OP_COUNT] = SX_WRITES_TABLE;
__constant__ DepTable c_dep = make_dep_table();
__constant__ uint8_t c_vec[32][8] = SX_MID_VECTORS;

static const uint8_t h_reads[OP_COUNT] = SX_READS_TABLE;
static const uint8_t h_writes[OP_COUNT] = SX_WRITES_TABLE;
static const DepTable h_dep = make_dep_table();
static const uint8_t h_vec[32][8] = SX_MID_VECTORS;

constexpr int kWave = 64;
constexpr int kChunkLog = 14;           // states per exhaustive chunk
constexpr uint32_t kRankNone = 0xFFFFFFFFu;
constexpr uint32_t kQcBlk = 8;          // targets per QuickCheck hash block (one scalar load)
constexpr int kQcQueue = 512;           // LDS survivor queue entries per wave
// QuickCheck survivors go to 8 regions of the survivor buffer, each with its
// own counter (block b uses region b mod 8): a queue flush reserves its
// output with an atomic, and atomics on one address serialise, which bound
// the kernel where hits are dense.  MidCheck reads the regions as one list.
constexpr uint32_t kQcRegions = 8;
struct SurvRegions {
  uint32_t pre[kQcRegions + 1];  // region r holds list entries [pre[r], pre[r + 1])
  uint32_t cap;                  // entries per region
};
// The Mid survivor list in regions with a counter each, 256 bytes apart (a
// wave flushes into region wave mod kMidRegions: one counter for all flushes
// serialised the joined QuickCheck + MidCheck at --dead-flags 0xFF, 470M
// survivors); list entry i is in region r = the last with pre[r] <= i, at
// r * cap + i - pre[r].  A single list is region 0 with pre[1..] = its length.
constexpr uint32_t kMidRegions = 16, kMidCntStride = 64;
struct MidList {
  uint32_t pre[kMidRegions + 1];
  uint32_t cap;
  __device__ __forceinline__ uint64_t at(uint32_t i) const {
    uint32_t r = 0;
#pragma unroll
    for (uint32_t k = 1; k < kMidRegions; k++) r += i >= pre[k] ? 1u : 0u;
    return (uint64_t)r * cap + (i - pre[r]);
  }
  __device__ __forceinline__ uint32_t size() const { return pre[kMidRegions]; }
};
constexpr int kEvalShards = 1024;       // shards of the evaluated-states counter
constexpr size_t kPinBytes = 64 << 10;  // pinned read-back staging per handle
constexpr uint64_t kCopyPiece = 1u << 16;  // rules per piece of the copy to the host


// ------------------------------------------------------ device helpers --

__device__ __forceinline__ uint32_t uni(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}
__device__ __forceinline__ uint64_t uni64(uint64_t v) {
  return (uint64_t)uni((uint32_t)v) | ((uint64_t)uni((uint32_t)(v >> 32)) << 32);
}

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & (kWave - 1); }

// Lanes whose predicate holds, as the compare's lane mask (no round trip
// through a VGPR).
__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

__device__ __forceinline__ uint32_t mbcnt(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
}

__device__ __forceinline__ uint64_t wave_sum64(uint64_t v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__device__ __forceinline__ void run_seq_uniform(State &st, const uint32_t *w, int n) {
  for (int i = 0; i < n; i++) {
    uint32_t word = uni(w[i]);
    step(st, word & 0xFFFFu, word >> 16);
  }
}

__device__ __forceinline__ State load_vec(uint32_t v) {
  State s;
  s.a = c_vec[v][0]; s.x = c_vec[v][1]; s.y = c_vec[v][2]; s.p = c_vec[v][3];
  s.s = c_vec[v][4]; s.m = c_vec[v][5]; s.s0 = c_vec[v][6]; s.s1 = c_vec[v][7];
  return s;
}

// ---------------------------------------------------------------- kernels --

// Ordered compaction of the pruned length-2 targets (search.go:56-75):
// pass 1 counts survivors per 256-pair block, pass 2 writes them in Go order.
__global__ __launch_bounds__(256) void k_gen_count(uint32_t *block_counts) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  bool keep = false;
  if (i < (uint32_t)kNumInsts * kNumInsts) {
    uint32_t seq[2] = {go_inst(i / kNumInsts), go_inst(i % kNumInsts)};
    keep = !should_prune(seq, 2, c_reads, c_writes);
  }
  __shared__ uint32_t cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  uint64_t m = __ballot(keep);
  if (lane_id() == 0) atomicAdd(&cnt, (uint32_t)__popcll(m));
  __syncthreads();
  if (threadIdx.x == 0) block_counts[blockIdx.x] = cnt;
}

__global__ __launch_bounds__(256) void k_gen_write(const uint64_t *block_offsets, uint32_t *out) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  bool keep = false;
  uint32_t seq[2] = {0, 0};
  if (i < (uint32_t)kNumInsts * kNumInsts) {
    seq[0] = go_inst(i / kNumInsts);
    seq[1] = go_inst(i % kNumInsts);
    keep = !should_prune(seq, 2, c_reads, c_writes);
  }
  __shared__ uint32_t wave_cnt[4];
  uint64_t m = __ballot(keep);
  uint32_t w = threadIdx.x >> 6;
  if (lane_id() == 0) wave_cnt[w] = (uint32_t)__popcll(m);
  __syncthreads();
  uint64_t base = block_offsets[blockIdx.x];
  for (uint32_t k = 0; k < w; k++) base += wave_cnt[k];
  if (keep) {
    uint64_t o = base + mbcnt(m);
    out[2 * o] = seq[0];
    out[2 * o + 1] = seq[1];
  }
}

// Pruned length-3 targets of a range of raw Go-order triple indices
// [r0, r0 + n) (index = i0 * 2863^2 + i1 * 2863 + i2): the same two-pass
// ordered compaction, also recording each survivor's offset in the range.
__device__ __forceinline__ void triple_words(uint64_t idx, uint32_t seq[3]) {
  const uint64_t sq = (uint64_t)kNumInsts * kNumInsts;
  seq[0] = go_inst((uint32_t)(idx / sq));
  seq[1] = go_inst((uint32_t)((idx / kNumInsts) % kNumInsts));
  seq[2] = go_inst((uint32_t)(idx % kNumInsts));
}

__global__ __launch_bounds__(256) void k_gen3_count(uint64_t r0, uint32_t n, uint32_t *block_counts) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  bool keep = false;
  if (i < n) {
    uint32_t seq[3];
    triple_words(r0 + i, seq);
    keep = !should_prune(seq, 3, c_reads, c_writes);
  }
  __shared__ uint32_t cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  uint64_t m = __ballot(keep);
  if (lane_id() == 0) atomicAdd(&cnt, (uint32_t)__popcll(m));
  __syncthreads();
  if (threadIdx.x == 0) block_counts[blockIdx.x] = cnt;
}

__global__ __launch_bounds__(256) void k_gen3_write(uint64_t r0, uint32_t n, const uint64_t *block_offsets,
                                                    uint32_t *out, uint32_t *out_off) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  bool keep = false;
  uint32_t seq[3] = {0, 0, 0};
  if (i < n) {
    triple_words(r0 + i, seq);
    keep = !should_prune(seq, 3, c_reads, c_writes);
  }
  __shared__ uint32_t wave_cnt[4];
  uint64_t m = __ballot(keep);
  uint32_t w = threadIdx.x >> 6;
  if (lane_id() == 0) wave_cnt[w] = (uint32_t)__popcll(m);
  __syncthreads();
  uint64_t base = block_offsets[blockIdx.x];
  for (uint32_t k = 0; k < w; k++) base += wave_cnt[k];
  if (keep) {
    uint64_t o = base + mbcnt(m);
    out[3 * o] = seq[0];
    out[3 * o + 1] = seq[1];
    out[3 * o + 2] = seq[2];
    out_off[o] = i;
  }
}

// Exclusive scan of uint32 counts into uint64 offsets, reduce-then-scan over
// tiles of 4096 counts (256 threads x 16 consecutive counts, read as four
// 16-byte loads): k_tile_sums writes each tile's sum, k_scan_sums scans the
// sums in one block, k_tile_scan re-reads the tile and writes its offsets.
// The counts are read twice and the offsets written once.
constexpr uint32_t kScanItems = 16, kScanTile = 256 * kScanItems;

__device__ __forceinline__ void scan_load(const uint32_t *in, uint64_t n, uint64_t base, uint32_t (&v)[kScanItems]) {
  if (base + kScanItems <= n) {
    const uint4 *p = reinterpret_cast<const uint4 *>(in + base);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const uint4 x = p[q];
      v[4 * q] = x.x;
      v[4 * q + 1] = x.y;
      v[4 * q + 2] = x.z;
      v[4 * q + 3] = x.w;
    }
  } else {
#pragma unroll
    for (int i = 0; i < (int)kScanItems; i++) v[i] = base + i < n ? in[base + i] : 0u;
  }
}

// Exclusive prefix of x over the 256 threads of the block, and the block's
// total; it synchronises once (s_w: 4 entries, not reused by the caller).
__device__ __forceinline__ uint64_t block_scan256(uint64_t x, uint64_t *s_w, uint64_t *total) {
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  uint64_t inc = x;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint64_t y = __shfl_up(inc, d);
    if (lane >= (uint32_t)d) inc += y;
  }
  if (lane == 63) s_w[w] = inc;
  __syncthreads();
  uint64_t pre = 0, tot = 0;
#pragma unroll
  for (uint32_t k = 0; k < 4; k++) {
    pre += k < w ? s_w[k] : 0ull;
    tot += s_w[k];
  }
  *total = tot;
  return pre + inc - x;
}

// Flags: count the nonzero entries instead of adding them.
__global__ __launch_bounds__(256) void k_tile_sums(const uint32_t *in, uint64_t n, uint64_t *sums, int flags) {
  __shared__ uint64_t s_w[4];
  uint32_t v[kScanItems];
  scan_load(in, n, (uint64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems, v);
  uint64_t t = 0;
#pragma unroll
  for (int i = 0; i < (int)kScanItems; i++) t += flags ? (uint64_t)(v[i] != 0u) : (uint64_t)v[i];
  uint64_t tot = 0;
  (void)block_scan256(t, s_w, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void k_tile_scan(const uint32_t *in, uint64_t n, const uint64_t *sums,
                                                   uint64_t *out) {
  __shared__ uint64_t s_w[4];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  scan_load(in, n, base, v);
  uint64_t t = 0;
#pragma unroll
  for (int i = 0; i < (int)kScanItems; i++) t += v[i];
  uint64_t tot = 0;
  uint64_t run = block_scan256(t, s_w, &tot) + sums[blockIdx.x];
  uint64_t o[kScanItems];
#pragma unroll
  for (int i = 0; i < (int)kScanItems; i++) {
    o[i] = run;
    run += v[i];
  }
  if (base + kScanItems <= n) {
    ulonglong2 *q = reinterpret_cast<ulonglong2 *>(out + base);
#pragma unroll
    for (int i = 0; i < (int)kScanItems / 2; i++) q[i] = make_ulonglong2(o[2 * i], o[2 * i + 1]);
  } else {
#pragma unroll
    for (int i = 0; i < (int)kScanItems; i++)
      if (base + i < n) out[base + i] = o[i];
  }
}

__global__ __launch_bounds__(1024) void k_scan_sums(uint64_t *sums, uint64_t nb, uint64_t *total) {
  __shared__ uint64_t s[1024];
  __shared__ uint64_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (uint64_t base = 0; base < nb; base += 1024) {
    uint64_t i = base + threadIdx.x;
    uint64_t v = i < nb ? sums[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      uint64_t t = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nb) sums[i] = s[threadIdx.x] - v + carry;
    __syncthreads();
    if (threadIdx.x == 1023) carry += s[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

// Filter hashes.  qc_mix(lo, hi, v) hashes one vector's output (64 bits and
// the vector index) to 32 bits with two rounds of the murmur3 finalizer,
// which spreads every input bit over the whole word; sequences are hashed by
// XOR over vectors.  (A single multiply leaves a top-bit difference in the
// top bit only, and two such differences cancel in the XOR: that version
// let through 13% more MidCheck pairs on the reference's sample.)  Equal
// outputs give equal hashes, so comparing hashes drops no pair whose
// outputs agree; a collision only costs the exact check that follows.
__device__ __forceinline__ uint32_t fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  return h ^ (h >> 16);
}
__device__ __forceinline__ uint32_t qc_mix(uint32_t lo_m, uint32_t hi, uint32_t v) {
  return fmix32(lo_m + fmix32(hi ^ (v * 0x9E3779B1u)));
}

// XOR of x over each 16-lane row, in every lane of the row: DPP moves
// (quad_perm [1,0,3,2], [2,3,0,1], row_half_mirror, row_mirror) at full VALU
// rate, no LDS round trip.  quad_only stops after the two quad steps.
template <bool QuadOnly = false>
__device__ __forceinline__ uint32_t row_xor(uint32_t x) {
  x ^= (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0xB1, 0xF, 0xF, false);
  x ^= (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x4E, 0xF, 0xF, false);
  if (QuadOnly) return x;
  x ^= (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x141, 0xF, 0xF, false);
  x ^= (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x140, 0xF, 0xF, false);
  return x;
}

// Outputs of sequences on the 32 Mid vectors.  A wave takes kFpPerWave
// sequences, two per pass: lanes 0..31 run sequence q on vector (lane & 31)
// and lanes 32..63 sequence q + 1.  The opcodes are wave-uniform (one scalar
// dispatch per instruction); the immediates are per lane, so the two halves
// share a pass whenever the sequences differ only in immediates (neighbours
// in the enumeration order nearly always do), and otherwise take a pass each
// with the other half idle.  When every sequence of the wave has the same
// words but the last (runs of targets with one first instruction), that
// prefix runs once per wave and the passes run only the last instruction.
// The words are fetched at once, one per lane; the per-sequence checks (an
// unknown opcode, same opcodes as the next sequence, same prefix) are lane
// work and ballots, since the scalar unit, shared by the CU's four SIMDs, is
// this kernel's bound.  fp[i * 8 + v] = (lo, hi) on QuickCheck vector v < 8,
// with P masked by 0xCF; mh[i] = (hash of the outputs on Mid vectors 8..31
// with P masked by 0xCF, the same with the lo words ANDed with lo_mask):
// MidCheck compares these instead of 24 vectors (equal outputs give equal
// hashes, so the filter drops no pair the exact check would keep); th[i]
// (when given) = the QuickCheck filter hash under lo_mask, densely, so that
// eight consecutive targets are one 32-byte scalar load in QuickCheck.
// lo_mask is the QuickCheck compare mask, 0x00FFFFFF | pm_masked << 24.  The
// hashes are XOR-reduced per 16-lane row by DPP, gathered into lane q of the
// wave and stored once per wave.  Lane k < kFpPerWave computes bytes, cycles
// and (when prune is given) the pruner flag of the wave's k-th sequence.
constexpr uint32_t kFpPerWave = 16;
constexpr uint32_t kFpVecs = 8;  // vectors kept per sequence: the QuickCheck ones
static_assert(kFpPerWave * 4 <= kWave, "sequence words (len <= 4) must fit one per lane");
static_assert(kFpPerWave % 2 == 0 && kFpPerWave <= 32, "sequences go two per pass; masks are 32-bit");
// With a list (fp_batch's `list`), the wave's sequences are list[first ..
// first + cnt) instead of first .. first + cnt; only mh is written (fp, th,
// bytes, cycles and the pruner flag are null).
__device__ __forceinline__ void fp_batch(const uint32_t *words, const uint32_t *list, uint32_t len, uint64_t first,
                                         uint64_t n, int bytes_model, uint2 *fp, uint2 *mh, uint32_t *th,
                                         uint32_t lo_mask, uint8_t *bytes, uint8_t *cyc, uint8_t *prune,
                                         uint32_t *bad_op) {
  const uint32_t lane = lane_id(), v = lane & 31, half = lane >> 5;
  const uint32_t cnt = (uint32_t)(first + kFpPerWave < n ? kFpPerWave : n - first);
  uint32_t wv = 0;  // len <= 4
  uint64_t my_seq = 0;  // lane q < cnt: the index of sequence q
  if (lane < cnt) my_seq = list ? list[first + lane] : first + lane;
  if (list && len) {
    const uint32_t j = lane / len;
    const uint64_t sj = (uint64_t)__shfl((int)(uint32_t)my_seq, (int)j);
    if (lane < cnt * len) wv = words[sj * len + (lane - j * len)];
  } else if (lane < cnt * len) {
    wv = words[first * len + lane];
  }
  // lane q < cnt: flags of sequence q (the other lanes read sequence 0 and
  // are masked out of the ballots)
  const uint32_t cmask = cnt >= 32 ? 0xFFFFFFFFu : (1u << cnt) - 1u;
  uint32_t bad_seqs, same_next;
  bool share;
  {
    const uint32_t q = lane < cnt ? lane : 0u, qn = q + 1 < cnt ? q + 1 : q;
    bool l_bad = false, l_same = q + 1 < cnt, l_pre = true;
    for (uint32_t k = 0; k < len; k++) {
      const uint32_t w = (uint32_t)__shfl((int)wv, (int)(q * len + k));
      const uint32_t wn = (uint32_t)__shfl((int)wv, (int)(qn * len + k));
      l_bad |= (w & 0xFFFFu) >= OP_COUNT;
      l_same &= (w & 0xFFFFu) == (wn & 0xFFFFu);
      if (k + 1 < len) l_pre &= w == __builtin_amdgcn_readlane(wv, k);
    }
    bad_seqs = (uint32_t)ballot(l_bad) & cmask;
    same_next = (uint32_t)ballot(l_same) & cmask;
    share = len > 1 && ((uint32_t)ballot(l_pre) & cmask) == cmask;
  }
  const bool two_masks = (lo_mask >> 24) != kPMask;  // else the masked hash is the exact one
  State base = load_vec(v);
  uint32_t k0 = 0;  // first word the passes run
  if (share) {
    for (uint32_t k = 0; k + 1 < len; k++) {
      const uint32_t w = __builtin_amdgcn_readlane(wv, k);
      step(base, w & 0xFFFFu, w >> 16);  // an unknown opcode marks every sequence bad
    }
    k0 = len - 1;
  }
  uint32_t g_hx = 0, g_hm = 0, g_th = 0;  // lane q: the hashes of sequence q
  for (uint32_t q = 0; q < cnt; q += 2) {
    const bool pair = q + 1 < cnt;
    const bool same = (same_next >> q) & 1u;  // implies pair
    // pass 0 runs q (and q + 1 when same), pass 1 runs q + 1 alone
    const uint32_t npass = pair && !same ? 2u : 1u;
    for (uint32_t ps = 0; ps < npass; ps++) {
      const uint32_t s_lo = same ? q : q + ps;  // sequence of half 0 (and of half 1 when same)
      const uint32_t sq = same ? q + half : q + ps;
      const bool run = (same ? true : (pair ? half == ps : half == 0)) && !((bad_seqs >> sq) & 1u);
      uint32_t hv = 0, hx = 0, hm = 0;
      if (run) {
        State st = base;
        for (uint32_t k = k0; k < len; k++) {
          const uint32_t w = __builtin_amdgcn_readlane(wv, s_lo * len + k);
          const uint32_t wb = same ? __builtin_amdgcn_readlane(wv, (s_lo + 1) * len + k) : w;
          step(st, w & 0xFFFFu, (half ? wb : w) >> 16);
        }
        st.p &= kPMask;
        const uint2 o = make_uint2(pack_lo(st), pack_hi(st));
        if (fp && v < kFpVecs) fp[(first + sq) * kFpVecs + v] = o;
        // one mix per lane: QuickCheck vectors 0..3 under the compare mask,
        // Mid vectors 8..31 exact (the masked Mid hash is a second mix, only
        // when the masks differ)
        const uint32_t h = qc_mix(v < 4 ? o.x & lo_mask : o.x, o.y, v);
        hv = v < 4 ? h : 0u;
        hx = v >= kFpVecs ? h : 0u;
        if (two_masks && v >= kFpVecs) hm = qc_mix(o.x & lo_mask, o.y, v);
      }
      // full wave again: reduce per row, rows 0+1 are half 0, rows 2+3 half 1
      hv = row_xor<true>(hv);
      hx = row_xor(hx);
      const uint32_t tha = __builtin_amdgcn_readlane(hv, 0), thb = __builtin_amdgcn_readlane(hv, 32);
      const uint32_t hxa = __builtin_amdgcn_readlane(hx, 0) ^ __builtin_amdgcn_readlane(hx, 16);
      const uint32_t hxb = __builtin_amdgcn_readlane(hx, 32) ^ __builtin_amdgcn_readlane(hx, 48);
      uint32_t hma = hxa, hmb = hxb;
      if (two_masks) {
        hm = row_xor(hm);
        hma = __builtin_amdgcn_readlane(hm, 0) ^ __builtin_amdgcn_readlane(hm, 16);
        hmb = __builtin_amdgcn_readlane(hm, 32) ^ __builtin_amdgcn_readlane(hm, 48);
      }
      if (same || ps == 0) {  // half 0 ran sequence q
        const bool mine = lane == q;
        g_th = mine ? tha : g_th;
        g_hx = mine ? hxa : g_hx;
        g_hm = mine ? hma : g_hm;
      }
      if (pair && (same || ps == 1)) {  // half 1 ran sequence q + 1
        const bool mine = lane == q + 1;
        g_th = mine ? thb : g_th;
        g_hx = mine ? hxb : g_hx;
        g_hm = mine ? hmb : g_hm;
      }
    }
  }
  if (lane < cnt && !((bad_seqs >> lane) & 1u)) {
    mh[my_seq] = make_uint2(g_hx, g_hm);
    if (th) th[first + lane] = g_th;
  }
  const bool bad_any = bad_seqs != 0;
  if (bad_any && lane == 0) atomicOr(bad_op, 1u);
  const uint64_t seq = first + lane;
  if (bytes && lane < cnt) {
    uint32_t w[4] = {0, 0, 0, 0};
    bool bad = false;
    for (uint32_t k = 0; k < len; k++) {
      w[k] = words[seq * len + k];
      bad |= (w[k] & 0xFFFFu) >= OP_COUNT;
    }
    if (!bad) {
      int b = 0, c = 0;
      for (uint32_t k = 0; k < len; k++) {
        b += byte_size(w[k] & 0xFFFFu, bytes_model);
        c += cycles(w[k] & 0xFFFFu);
      }
      bytes[seq] = (uint8_t)b;
      cyc[seq] = (uint8_t)c;
      if (prune) prune[seq] = should_prune(w, (int)len, c_reads, c_writes) ? 1 : 0;
    }
  }
}
__global__ __launch_bounds__(256) void k_fingerprints(const uint32_t *words, uint32_t len,
                                                      uint64_t n, int bytes_model, uint2 *fp, uint2 *mh,
                                                      uint32_t *th, uint32_t lo_mask, uint8_t *bytes, uint8_t *cyc,
                                                      uint8_t *prune,
                                                      uint32_t *bad_op) {
  const uint64_t first = ((uint64_t)blockIdx.x * 4 + uni(threadIdx.x >> 6)) * kFpPerWave;  // wave-uniform
  if (first >= n) return;
  fp_batch(words, nullptr, len, first, n, bytes_model, fp, mh, th, lo_mask, bytes, cyc, prune, bad_op);
}
// The Mid hashes of the sequences list[0 .. *list_n), waves striding over the
// list (its length is known on the device only).
__global__ __launch_bounds__(256) void k_fp_mid_list(const uint32_t *words, const uint32_t *list,
                                                     const uint32_t *list_n, uint32_t len, uint2 *mh, uint32_t lo_mask,
                                                     uint32_t *bad_op) {
  const uint64_t n = uni(*list_n);
  const uint64_t stride = (uint64_t)gridDim.x * 4 * kFpPerWave;
  for (uint64_t first = ((uint64_t)blockIdx.x * 4 + uni(threadIdx.x >> 6)) * kFpPerWave; first < n;
       first += stride)
    fp_batch(words, list, len, first, n, 0, nullptr, mh, nullptr, lo_mask, nullptr, nullptr, nullptr, bad_op);
}

// QuickCheck.  Grid: x = 256-candidate group, y = target tile.
// Survivor entry: (target, slot | full << 31), where full = the exact
// (0xCF) compare holds; every entry satisfies the compare under pm_masked.
// One QuickCheck work item: 256 candidate slots (lane = candidate, its 8 QC
// outputs in registers) against the targets [t0, t1).  Survivors go to the
// calling wave's LDS queue (fill qn, wave-uniform) and are flushed with one
// atomic per batch.
struct QcQueue {
  uint2 *q;       // this wave's LDS queue
  uint32_t qn;
  uint2 *surv;
  uint32_t cap;
  uint32_t *count;

  __device__ __forceinline__ void flush(uint32_t lane) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(count, qn);
    base = uni(base);
    for (uint32_t i = lane; i < qn; i += kWave) {
      uint32_t o = base + i;
      if (o < cap) surv[o] = q[i];
    }
    __builtin_amdgcn_wave_barrier();
    qn = 0;
  }
  // End of a block: the four waves' queues in one reservation (one atomic per
  // block instead of one per wave; atomics on one address serialise, and
  // nearly every wave ends with survivors queued when hits are dense).  Every
  // thread of the block calls it.
  __device__ __forceinline__ void flush_block(uint32_t lane, uint32_t *s_cnt, uint32_t *s_base) {
    const uint32_t w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) s_cnt[w] = qn;
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint32_t tot = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
      *s_base = tot ? atomicAdd(count, tot) : 0u;
    }
    __syncthreads();
    uint32_t base = *s_base;
    for (uint32_t k = 0; k < w; k++) base += s_cnt[k];
    for (uint32_t i = lane; i < qn; i += kWave) {
      const uint32_t o = base + i;
      if (o < cap) surv[o] = q[i];
    }
    qn = 0;
  }
};
static_assert(kWave * 4 == 256, "QuickCheck blocks are four waves");

__device__ __forceinline__ void qc_item(const uint32_t *__restrict__ th, const uint2 *__restrict__ tfp, uint32_t nt,
                                        uint32_t t0, uint32_t t1, const uint2 *__restrict__ cfp, uint32_t nc,
                                        uint32_t slot, uint32_t pm_masked, QcQueue &Q) {
  const uint32_t lane = lane_id();
  const bool valid = slot < nc;
  uint32_t clo[8], chi[8];
#pragma unroll
  for (int v = 0; v < 8; v++) {
    uint2 f = valid ? cfp[(uint64_t)slot * kFpVecs + v] : make_uint2(0, 0);
    clo[v] = f.x;
    chi[v] = f.y;
  }
  const uint32_t mlo = 0x00FFFFFFu | (pm_masked << 24);
  const uint32_t flo = 0x00FFFFFFu | (kPMask << 24);
  // the candidate's filter hash, as the targets' plane holds it
  const uint32_t hc = qc_mix(clo[0] & mlo, chi[0], 0) ^ qc_mix(clo[1] & mlo, chi[1], 1) ^
                      qc_mix(clo[2] & mlo, chi[2], 2) ^ qc_mix(clo[3] & mlo, chi[3], 3);
  const uint64_t vmask = ballot(valid);
  // 8 consecutive target hashes, one 32-byte scalar load (the plane has
  // kQcBlk entries of slack past nt, so no clamping is needed)
  struct alignas(32) ThBlk {
    uint32_t h[kQcBlk];
  };
  // target tb + u against the wave's candidates: hash filter, then all 8
  // vectors exactly under both masks (the differences ORed together first:
  // one 3-input VALU op per dword)
  auto check = [&](const ThBlk &blk, uint32_t tb, uint32_t u) {
    const uint64_t m = ballot(hc == blk.h[u]) & vmask;
    if (m == 0) return;
    const uint2 *tf = tfp + (uint64_t)(tb + u) * kFpVecs;
    uint32_t acc_lo = 0, acc_hi = 0;
#pragma unroll
    for (int v = 0; v < 8; v++) {
      const uint2 f = tf[v];
      acc_lo |= clo[v] ^ f.x;
      acc_hi |= chi[v] ^ f.y;
    }
    const bool okm = valid && ((acc_lo & mlo) | acc_hi) == 0;
    const bool okf = valid && ((acc_lo & flo) | acc_hi) == 0;
    const uint64_t hm = __ballot(okm);
    if (hm == 0) return;
    if (okm) Q.q[Q.qn + mbcnt(hm)] = make_uint2(tb + u, slot | (okf ? 0x80000000u : 0u));
    Q.qn += (uint32_t)__popcll(hm);
    if (Q.qn > kQcQueue - kWave) Q.flush(lane);
  };
  // the minimum of the differences is 0 iff some target's hash matches: v_xor
  // and a v_min3 tree on the vector unit, one ballot and one branch per 32
  // targets (the scalar unit, shared by the CU's four SIMDs, was the bound
  // when each compare was its own scalar mask); four blocks' loads are in
  // flight together
  auto dmin_of = [&](const ThBlk &blk) {
    uint32_t d = hc ^ blk.h[0];
#pragma unroll
    for (uint32_t j = 1; j < kQcBlk; j++) d = min(d, hc ^ blk.h[j]);
    return d;
  };
  uint32_t t = t0;  // a multiple of kQcBlk
  for (; t + 4 * kQcBlk <= t1; t += 4 * kQcBlk) {
    ThBlk b[4];
#pragma unroll
    for (int k = 0; k < 4; k++) b[k] = *reinterpret_cast<const ThBlk *>(th + t + k * kQcBlk);
    uint32_t d[4];
#pragma unroll
    for (int k = 0; k < 4; k++) d[k] = dmin_of(b[k]);
    if ((ballot(min(min(d[0], d[1]), min(d[2], d[3])) == 0) & vmask) == 0) continue;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      if (ballot(d[k] == 0) & vmask) {
#pragma unroll
        for (uint32_t u = 0; u < kQcBlk; u++) check(b[k], t + k * kQcBlk, u);
      }
    }
  }
  for (; t < t1; t += kQcBlk) {
    const uint32_t nb = min(kQcBlk, t1 - t);
    const ThBlk blk = *reinterpret_cast<const ThBlk *>(th + t);
#pragma unroll
    for (uint32_t u = 0; u < kQcBlk; u++) {
      if (u >= nb) break;
      check(blk, t, u);
    }
  }
}

// Grid form: block (x, y) = candidate slots [256 x, 256 x + 256) against
// target tile y.
__global__ __launch_bounds__(256) void k_quickcheck(const uint32_t *__restrict__ th,
                                                    const uint2 *__restrict__ tfp, uint32_t nt,
                                                    uint32_t tile, const uint2 *__restrict__ cfp,
                                                    uint32_t nc, uint32_t pm_masked,
                                                    uint2 *__restrict__ surv, uint32_t region_cap,
                                                    uint32_t *__restrict__ surv_count) {
  __shared__ uint2 queue[4][kQcQueue];
  __shared__ uint32_t s_cnt[4], s_base;
  const uint32_t reg = (blockIdx.x + blockIdx.y * gridDim.x) & (kQcRegions - 1);
  QcQueue Q{queue[threadIdx.x >> 6], 0, surv + (uint64_t)reg * region_cap, region_cap, surv_count + reg};
  const uint32_t t0 = blockIdx.y * tile;
  qc_item(th, tfp, nt, t0, min(nt, t0 + tile), cfp, nc, blockIdx.x * 256 + threadIdx.x, pm_masked, Q);
  Q.flush_block(lane_id(), s_cnt, &s_base);
}

// QuickCheck as a hash join.  A pair can pass only if the target's and the
// candidate's filter hashes (vectors 0-3 under the masked compare) are equal,
// so instead of comparing every pair's hashes the candidates are sorted by
// hash once per enumeration, with a directory of 2^dk buckets on the top
// hash bits (dir[b] = first sorted position of bucket b, dir[2^dk] = nc),
// and each target reads its bucket.  A wave takes 64 targets; the bucket
// sizes are prefix-summed across the wave and the wave walks the flattened
// (target, bucket entry) items 64 at a time, so a target with hundreds of
// equal-hash candidates spreads over the lanes.  An entry with an equal hash
// gets the exact 8-vector compare under both masks, as in qc_item, and the
// survivors leave through the same queue into the same regions: the survivor
// set is exactly the all-pairs kernel's.
__global__ __launch_bounds__(256) void k_cand_filter_hash(const uint2 *__restrict__ cfp, uint64_t nc, uint32_t lo_mask,
                                                          uint32_t *key, uint32_t *val) {
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= nc) return;
  uint32_t h = 0;
#pragma unroll
  for (uint32_t v = 0; v < 4; v++) {
    const uint2 f = cfp[i * kFpVecs + v];
    h ^= qc_mix(f.x & lo_mask, f.y, v);
  }
  key[i] = h;
  val[i] = (uint32_t)i;
}
__global__ __launch_bounds__(256) void k_hash_dir(const uint32_t *__restrict__ shash, uint32_t nc, uint32_t dk,
                                                  uint32_t *dir) {
  const uint32_t b = blockIdx.x * 256 + threadIdx.x;
  if (b > (1u << dk)) return;
  if (b == (1u << dk)) {
    dir[b] = nc;
    return;
  }
  const uint32_t k = b << (32 - dk);  // first hash of the bucket
  uint32_t lo = 0, hi = nc;           // lower bound of k
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (shash[mid] < k) lo = mid + 1;
    else hi = mid;
  }
  dir[b] = lo;
}

// Bucket directory of the sorted 64-bit class keys on their top dk bits:
// dir[b] = the first class whose key is in bucket b or later, dir[2^dk] =
// the class count (known on the device only).
__global__ __launch_bounds__(256) void k_class_dir(const unsigned long long *__restrict__ ckey,
                                                   const uint64_t *__restrict__ ncls_p, uint32_t dk, uint32_t *dir) {
  const uint32_t b = blockIdx.x * 256 + threadIdx.x;
  if (b > (1u << dk)) return;
  const uint32_t ncls = (uint32_t)*ncls_p;
  if (b == (1u << dk)) {
    dir[b] = ncls;
    return;
  }
  const unsigned long long k = (unsigned long long)b << (64 - dk);  // first key of the bucket
  uint32_t lo = 0, hi = ncls;
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (ckey[mid] < k) lo = mid + 1;
    else hi = mid;
  }
  dir[b] = lo;
}

__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t x, uint32_t lane) {
  uint32_t s = x;
#pragma unroll
  for (uint32_t o = 1; o < kWave; o <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)s, o);
    if (lane >= o) s += y;
  }
  return s - x;
}

__global__ __launch_bounds__(256) void k_qc_join(const uint32_t *__restrict__ th, const uint2 *__restrict__ tfp,
                                                 uint32_t nt, const uint32_t *__restrict__ shash,
                                                 const uint32_t *__restrict__ sslot, const uint32_t *__restrict__ dir,
                                                 uint32_t dk, const uint2 *__restrict__ cfp, uint32_t pm_masked,
                                                 uint2 *__restrict__ surv, uint32_t region_cap,
                                                 uint32_t *__restrict__ surv_count) {
  __shared__ uint2 queue[4][kQcQueue];
  __shared__ uint32_t s_cnt[4], s_base;
  __shared__ uint32_t s_pre[4][kWave], s_lo[4][kWave], s_h[4][kWave];
  const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
  QcQueue Q{queue[w], 0, surv + (uint64_t)(blockIdx.x & (kQcRegions - 1)) * region_cap, region_cap,
            surv_count + (blockIdx.x & (kQcRegions - 1))};
  const uint32_t t0 = (blockIdx.x * 4 + w) * kWave;  // the wave's first target
  const uint32_t t = t0 + lane;
  const bool have = t < nt;
  const uint32_t h = have ? th[t] : 0u;
  const uint32_t b = h >> (32 - dk);
  const uint32_t lo = have ? dir[b] : 0u, c = have ? dir[b + 1] - lo : 0u;
  const uint32_t pre = wave_excl_scan(c, lane);
  const uint32_t total = __builtin_amdgcn_readlane(pre + c, kWave - 1);
  s_pre[w][lane] = pre;
  s_lo[w][lane] = lo;
  s_h[w][lane] = h;
  __builtin_amdgcn_wave_barrier();
  const uint32_t mlo = 0x00FFFFFFu | (pm_masked << 24);
  const uint32_t flo = 0x00FFFFFFu | (kPMask << 24);
  for (uint32_t base = 0; base < total; base += kWave) {
    const uint32_t i = base + lane;
    bool okm = false, okf = false;
    uint32_t tt = 0, slot = 0;
    if (i < total) {
      // the owning target: the last lane whose prefix is <= i (empty buckets
      // share their prefix with the next lane, which then owns the item)
      uint32_t o = 0;
#pragma unroll
      for (uint32_t step = kWave / 2; step >= 1; step >>= 1)
        if (s_pre[w][o + step] <= i) o += step;
      const uint32_t e = s_lo[w][o] + (i - s_pre[w][o]);
      if (shash[e] == s_h[w][o]) {
        tt = t0 + o;
        slot = sslot[e];
        const uint2 *tf = tfp + (uint64_t)tt * kFpVecs;
        const uint2 *cf = cfp + (uint64_t)slot * kFpVecs;
        uint32_t acc_lo = 0, acc_hi = 0;
#pragma unroll
        for (int v = 0; v < 8; v++) {
          const uint2 x = tf[v], y = cf[v];
          acc_lo |= x.x ^ y.x;
          acc_hi |= x.y ^ y.y;
        }
        okm = ((acc_lo & mlo) | acc_hi) == 0;
        okf = ((acc_lo & flo) | acc_hi) == 0;
      }
    }
    const uint64_t hm = ballot(okm);
    if (hm) {
      if (okm) Q.q[Q.qn + mbcnt(hm)] = make_uint2(tt, slot | (okf ? 0x80000000u : 0u));
      Q.qn += (uint32_t)__popcll(hm);
      if (Q.qn > kQcQueue - kWave) Q.flush(lane);
    }
  }
  Q.flush_block(lane, s_cnt, &s_base);
}

// Persistent form: a fixed grid pulls work items (candidate block, target
// tile) from 8 device queues, one per XCD (workgroups are dispatched to the
// XCDs round-robin, so block b runs on XCD b mod 8).  Queue q owns the target
// tiles q, q + 8, ... so a tile's target data is read through one XCD's L2;
// a block drains its own queue first, then the others.  Every block leaves
// once all 8 queues are drained.
__global__ __launch_bounds__(256) void k_qc_persistent(const uint32_t *__restrict__ th,
                                                       const uint2 *__restrict__ tfp, uint32_t nt, uint32_t tile,
                                                       const uint2 *__restrict__ cfp, uint32_t nc,
                                                       uint32_t pm_masked, uint2 *__restrict__ surv,
                                                       uint32_t region_cap, uint32_t *__restrict__ surv_count,
                                                       uint32_t *__restrict__ heads, uint32_t ncb,
                                                       uint32_t ntiles) {
  __shared__ uint2 queue[4][kQcQueue];
  __shared__ uint32_t item_s, s_cnt[4], s_base;
  const uint32_t reg = blockIdx.x & (kQcRegions - 1);
  QcQueue Q{queue[threadIdx.x >> 6], 0, surv + (uint64_t)reg * region_cap, region_cap, surv_count + reg};
  const uint32_t home = blockIdx.x & 7u;
  for (uint32_t k = 0; k < 8; k++) {
    const uint32_t q = (home + k) & 7u;
    const uint32_t tiles_q = q < ntiles ? (ntiles - q + 7) / 8 : 0;
    const uint32_t nitems = tiles_q * ncb;
    for (;;) {
      __syncthreads();
      if (threadIdx.x == 0) item_s = nitems ? atomicAdd(&heads[q], 1u) : 0xFFFFFFFFu;
      __syncthreads();
      const uint32_t item = item_s;
      if (item >= nitems) break;
      const uint32_t tt = q + 8 * (item / ncb), cb = item % ncb;
      const uint32_t t0 = tt * tile;
      qc_item(th, tfp, nt, t0, min(nt, t0 + tile), cfp, nc, cb * 256 + threadIdx.x, pm_masked, Q);
    }
  }
  Q.flush_block(lane_id(), s_cnt, &s_base);
}

// Compaction kernels run 1,024-thread blocks and reserve their output with
// one atomic per block: atomics on one address serialise (about 9e7 per
// second here), so one per wave would cost more than the work.
constexpr int kBlk = 1024, kBlkWaves = kBlk / kWave;

// Per-wave counts -> this wave's exclusive prefix within the block and the
// block total.  Every thread of the block calls it; it synchronises twice.
__device__ __forceinline__ uint32_t block_prefix(uint32_t cnt, uint32_t *s, uint32_t *total) {
  const uint32_t w = threadIdx.x >> 6;
  __syncthreads();  // s is free
  if (lane_id() == 0) s[w] = cnt;
  __syncthreads();
  uint32_t pre = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < kBlkWaves; k++) {
    const uint32_t c = s[k];
    pre += (uint32_t)k < w ? c : 0u;
    tot += c;
  }
  *total = tot;
  return pre;
}

// dst[key] = min(dst[key], v) for every lane whose key is not kRankNone, with
// one atomic per run of equal keys in the wave: a suffix minimum over lanes of
// the same key (Hillis-Steele, 6 shuffle steps), then the first lane of each
// run issues the atomic.  Survivor lists keep a wave's hits on one target
// next to each other, and atomics on one address serialise, so at dense
// targets this cuts the atomics by the run length.  Every lane of the wave
// calls it.
__device__ __forceinline__ void wave_run_min(uint32_t key, uint32_t v, uint32_t *dst) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const uint32_t ok = __shfl_down(key, d), ov = __shfl_down(v, d);
    if (lane + d < (uint32_t)kWave && ok == key) v = min(v, ov);
  }
  const uint32_t pk = __shfl_up(key, 1);
  if (key != kRankNone && (lane == 0 || pk != key)) atomicMin(&dst[key], v);
}

// MidCheck and stage bookkeeping.  One lane per QuickCheck survivor.
// hits_all = 0 (Go): pass 1 if the exact compare held, pass 2 if only the
// masked one did (and dead != 0).  hits_all = 1 (CUDA tool): one masked pass.
// The 24 Mid vectors are compared through their hashes (tmh/cmh: exact and
// masked P), which keeps every pair whose outputs agree.  Output entry:
// (target, slot | pass << 30).  counters[0] += QuickCheck hits after the
// byte/prune filters and counters[1] += passes, both with one 64-bit atomic
// per block (counters[1] is also the output cursor).  cur0 (first hit, else
// null): the first round's minimum, cur0[target * 2 + pass - 1] = lowest rank
// among the entry's passing survivors (cur0 holds kRankNone on entry).
__global__ __launch_bounds__(kBlk) void k_midcheck(const uint2 *__restrict__ surv, const SurvRegions R, uint32_t ns,
                                                   const uint2 *__restrict__ tmh,
                                                   const uint8_t *__restrict__ tbytes,
                                                   const uint2 *__restrict__ cmh,
                                                   const uint8_t *__restrict__ cbytes,
                                                   const uint8_t *__restrict__ cprune,
                                                   int hits_all, int use_mid,
                                                   uint2 *__restrict__ out, uint32_t out_cap,
                                                   uint32_t *__restrict__ counters,
                                                   const uint32_t *__restrict__ rank, uint32_t *cur0) {
  __shared__ uint32_t s_hits[kBlkWaves], s_pass[kBlkWaves];
  __shared__ unsigned long long s_base;
  // Straight-line on purpose: the lane masks are computed with selects, not
  // with branches.  With uniform regions left unstructured (the build's
  // -structurizecfg-skip-uniform-regions), a flag set inside a divergent
  // branch and merged across the uniform use_mid branch came out as all
  // lanes of the outer region, which counted pruned candidates as hits.
  const uint32_t i = blockIdx.x * kBlk + threadIdx.x;
  const bool in = i < ns;
  // list entry i -> its region and place there (lanes past the list read entry 0)
  const uint32_t ii = in ? i : 0u;
  uint32_t base = 0, r = 0;
#pragma unroll
  for (uint32_t k = 1; k < kQcRegions; k++) {
    const bool after = ii >= R.pre[k];
    base = after ? R.pre[k] : base;
    r = after ? k : r;
  }
  const uint2 e = surv[(uint64_t)r * R.cap + (ii - base)];
  const uint32_t t = e.x, slot = e.y & 0x3FFFFFFFu;
  const bool full = (e.y >> 31) != 0;
  const bool qc_hit = in & (cbytes[slot] < tbytes[t]) & (cprune[slot] == 0);
  const bool exact = !hits_all && full;  // pass 1 compares P under 0xCF
  const uint32_t pass = (hits_all || full) ? 1u : 2u;
  const uint2 a = tmh[t], b = cmh[slot];
  const bool mid_ok = !use_mid || (exact ? a.x == b.x : a.y == b.y);
  const bool pass_ok = qc_hit & mid_ok;
  const uint64_t hq = __ballot(qc_hit), m = __ballot(pass_ok);
  uint32_t th = 0, tp = 0;
  block_prefix((uint32_t)__popcll(hq), s_hits, &th);
  const uint32_t pre = block_prefix((uint32_t)__popcll(m), s_pass, &tp);
  if (threadIdx.x == 0 && (th | tp))
    s_base = atomicAdd(reinterpret_cast<unsigned long long *>(counters), ((unsigned long long)tp << 32) | th);
  __syncthreads();
  if (pass_ok) {
    const uint32_t o = (uint32_t)(s_base >> 32) + pre + mbcnt(m);
    if (o < out_cap) out[o] = make_uint2(t, slot | (pass << 30));
  }
  if (cur0) wave_run_min(pass_ok ? t * 2 + (pass - 1) : kRankNone, pass_ok ? rank[slot] : 0u, cur0);
}

// QuickCheck and MidCheck as one join, with QuickCheck's counters from
// candidate classes (MidCheck on, either semantics).
//
// A pair survives MidCheck only if its outputs agree on the 8 QuickCheck
// vectors under the masked compare and its masked Mid hashes agree (the exact
// pass's tighter compares imply both).  So a join on jkey = f(hash of the 8
// masked QuickCheck outputs, masked Mid hash) reaches every MidCheck
// survivor, and each pair it reaches gets the exact QuickCheck compare and
// MidCheck's test as k_midcheck applies them: the survivor list is the same
// set, without the QuickCheck survivors (483M at --dead-flags 0xFF) ever
// being written.  QuickCheck's counters count pairs that pass QuickCheck
// whether or not they pass MidCheck; they come from classes of candidates
// with equal masked 8-vector outputs (sorted by a 64-bit key of those
// outputs; every member is compared with the class's first one, and a
// mismatch sends the enumeration back to the separate kernels): class c
// holds its size and its unpruned members per byte size, and a target whose
// outputs equal c's adds c's size and its unpruned members below the
// target's own size.
constexpr uint32_t kClsBins = 16;  // byte sizes 0..15 per class, then the class size
__device__ __forceinline__ void qc_keys(const uint2 (&o)[8], uint32_t lo_mask, uint32_t &k1, uint32_t &k2) {
  k1 = 0;
  k2 = 0;
#pragma unroll
  for (uint32_t v = 0; v < 8; v++) {
    const uint32_t x = o[v].x & lo_mask, y = o[v].y;
    k1 ^= qc_mix(x, y, v);
    k2 ^= fmix32(x * 0x9E3779B1u + fmix32(y ^ (v * 0x632BE5ABu + 0x7F4A7C15u)));
  }
}
__device__ __forceinline__ uint32_t join_key(uint32_t k1, uint32_t mh_masked) {
  return fmix32(k1 ^ fmix32(mh_masked + 0x9E3779B9u));
}
__device__ __forceinline__ void load_fp8(const uint2 *__restrict__ fp, uint64_t i, uint2 (&o)[8]) {
  const uint4 *f = reinterpret_cast<const uint4 *>(fp + i * kFpVecs);
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const uint4 q = f[k];
    o[2 * k] = make_uint2(q.x, q.y);
    o[2 * k + 1] = make_uint2(q.z, q.w);
  }
}
// differences of two outputs under the masked (okm) and exact (okf) compare
__device__ __forceinline__ void fp_compare(const uint2 (&a)[8], const uint2 (&b)[8], uint32_t mlo, bool &okm,
                                           bool &okf) {
  uint32_t acc_lo = 0, acc_hi = 0;
#pragma unroll
  for (int v = 0; v < 8; v++) {
    acc_lo |= a[v].x ^ b[v].x;
    acc_hi |= a[v].y ^ b[v].y;
  }
  okm = ((acc_lo & mlo) | acc_hi) == 0;
  okf = ((acc_lo & (0x00FFFFFFu | (kPMask << 24))) | acc_hi) == 0;
}
// The P bytes (under 0xCF) of the 8 QuickCheck outputs: two sequences that
// agree under the masked compare agree under the exact one iff these do.
__device__ __forceinline__ uint2 p_sig(const uint2 (&o)[8]) {
  return make_uint2((o[0].x >> 24) | ((o[1].x >> 24) << 8) | ((o[2].x >> 24) << 16) | (o[3].x & 0xFF000000u),
                    (o[4].x >> 24) | ((o[5].x >> 24) << 8) | ((o[6].x >> 24) << 16) | (o[7].x & 0xFF000000u));
}
constexpr uint32_t kNoClass = 0xFFFFFFFFu;

__global__ __launch_bounds__(256) void k_cand_keys(const uint2 *__restrict__ cfp, const uint2 *__restrict__ cmh,
                                                   uint64_t nc, uint32_t lo_mask, unsigned long long *ckey,
                                                   uint32_t *jkey, uint32_t *val1, uint32_t *val2, uint2 *psig) {
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= nc) return;
  uint2 o[8];
  load_fp8(cfp, i, o);
  psig[i] = p_sig(o);
  uint32_t k1, k2;
  qc_keys(o, lo_mask, k1, k2);
  ckey[i] = (unsigned long long)k1 | ((unsigned long long)k2 << 32);
  jkey[i] = join_key(k1, cmh[i].y);
  val1[i] = (uint32_t)i;
  val2[i] = (uint32_t)i;
}
__global__ __launch_bounds__(256) void k_class_heads(const unsigned long long *__restrict__ sk, uint32_t n,
                                                     uint32_t *head) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) head[i] = (i == 0 || sk[i] != sk[i - 1]) ? 1u : 0u;
}
// class ids from the heads' exclusive scan (ccls[slot] = its class); class
// tables zeroed beforehand
__global__ __launch_bounds__(256) void k_class_fill(const unsigned long long *__restrict__ sk,
                                                    const uint32_t *__restrict__ sslot, uint32_t n,
                                                    const uint32_t *__restrict__ head,
                                                    const uint64_t *__restrict__ hoff,
                                                    const uint8_t *__restrict__ cbytes,
                                                    const uint8_t *__restrict__ cprune, unsigned long long *ckey,
                                                    uint32_t *crep, uint32_t *ccnt, uint32_t *ccls) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint32_t cid = (uint32_t)hoff[i] + head[i] - 1u, slot = sslot[i];
  ccls[slot] = cid;
  if (head[i]) {
    ckey[cid] = sk[i];
    crep[cid] = slot;
  }
  atomicAdd(&ccnt[cid * (kClsBins + 1) + kClsBins], 1u);
  if (!cprune[slot]) atomicAdd(&ccnt[cid * (kClsBins + 1) + min((uint32_t)cbytes[slot], kClsBins - 1)], 1u);
}
__global__ __launch_bounds__(256) void k_class_verify(const uint32_t *__restrict__ sslot, uint32_t n,
                                                      const uint32_t *__restrict__ head,
                                                      const uint64_t *__restrict__ hoff,
                                                      const uint32_t *__restrict__ crep, const uint2 *__restrict__ cfp,
                                                      uint32_t lo_mask, uint32_t *bad) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint32_t cid = (uint32_t)hoff[i] + head[i] - 1u;
  uint2 a[8], b[8];
  load_fp8(cfp, sslot[i], a);
  load_fp8(cfp, crep[cid], b);
  bool okm, okf;
  fp_compare(a, b, lo_mask, okm, okf);
  if (!okm) atomicOr(bad, 1u);
}

constexpr uint32_t kQcCntShards = 64;
// QuickCheck's counters of a segment from the classes: sums[2 s] += raw
// survivors, sums[2 s + 1] += survivors after the byte filter and the pruner
// (shard s = block mod 64).  list (when given): the targets whose masked
// QuickCheck outputs equal some candidate's (every other target has no
// QuickCheck survivor), appended per block in target order at
// list[atomicAdd(list_n)] (one atomic per block: per-wave atomics on the one
// counter serialise).
constexpr uint32_t kQcCntChunks = 4;  // 256-target chunks per k_qc_count block
__global__ __launch_bounds__(256) void k_qc_count(const uint2 *__restrict__ tfp, const uint8_t *__restrict__ tbytes,
                                                  uint32_t nt, uint32_t lo_mask,
                                                  const unsigned long long *__restrict__ ckey,
                                                  const uint32_t *__restrict__ cdir, uint32_t cdk,
                                                  const uint32_t *__restrict__ crep, const uint32_t *__restrict__ ccnt,
                                                  const uint2 *__restrict__ cfp, unsigned long long *sums,
                                                  uint32_t *list, uint32_t *list_n, uint32_t *tcls) {
  // tcls[t] = the target's class (kNoClass without one)
  __shared__ unsigned long long s_raw[4], s_hit[4];
  __shared__ uint32_t s_in[kQcCntChunks][4], s_base;
  const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
  uint64_t raw = 0, hit = 0;
  uint64_t m[kQcCntChunks];  // per chunk: the wave's targets with a class
#pragma unroll
  for (uint32_t ch = 0; ch < kQcCntChunks; ch++) {
    const uint32_t t = (blockIdx.x * kQcCntChunks + ch) * 256 + threadIdx.x;
    bool in_class = false;
    if (t < nt) {
      uint2 o[8];
      load_fp8(tfp, t, o);
      uint32_t k1, k2;
      qc_keys(o, lo_mask, k1, k2);
      const unsigned long long k = (unsigned long long)k1 | ((unsigned long long)k2 << 32);
      // the class of key k: its bucket, then the bucket's few keys in order
      const uint32_t b = (uint32_t)(k >> (64 - cdk));
      uint32_t lo = cdir[b];
      const uint32_t hi = cdir[b + 1];
      while (lo < hi && ckey[lo] < k) lo++;
      if (lo < hi && ckey[lo] == k) {
        uint2 r[8];
        load_fp8(cfp, crep[lo], r);
        bool okm, okf;
        fp_compare(o, r, lo_mask, okm, okf);
        in_class = okm;
        if (okm) {
          const uint32_t *c = ccnt + (uint64_t)lo * (kClsBins + 1);
          raw += c[kClsBins];
          const uint32_t tb = min((uint32_t)tbytes[t], kClsBins);
          for (uint32_t b = 0; b < tb; b++) hit += c[b];
        }
      }
      tcls[t] = in_class ? lo : kNoClass;
    }
    m[ch] = ballot(in_class);
    if (lane == 0) s_in[ch][w] = (uint32_t)__popcll(m[ch]);
  }
  raw = wave_sum64(raw);
  hit = wave_sum64(hit);
  if (lane == 0) {
    s_raw[w] = raw;
    s_hit[w] = hit;
  }
  __syncthreads();
  if (list) {
    if (threadIdx.x == 0) {
      uint32_t tot = 0;
      for (uint32_t ch = 0; ch < kQcCntChunks; ch++) tot += s_in[ch][0] + s_in[ch][1] + s_in[ch][2] + s_in[ch][3];
      s_base = tot ? atomicAdd(list_n, tot) : 0u;
    }
    __syncthreads();
    uint32_t pre = s_base;
#pragma unroll
    for (uint32_t ch = 0; ch < kQcCntChunks; ch++) {
      uint32_t p = pre;
      for (uint32_t k = 0; k < w; k++) p += s_in[ch][k];
      if ((m[ch] >> lane) & 1u) list[p + mbcnt(m[ch])] = (blockIdx.x * kQcCntChunks + ch) * 256 + threadIdx.x;
      pre += s_in[ch][0] + s_in[ch][1] + s_in[ch][2] + s_in[ch][3];
    }
  }
  if (threadIdx.x == 0) {
    const uint64_t r = s_raw[0] + s_raw[1] + s_raw[2] + s_raw[3], hh = s_hit[0] + s_hit[1] + s_hit[2] + s_hit[3];
    const uint32_t sh = blockIdx.x & (kQcCntShards - 1);
    if (r) atomicAdd(&sums[2 * sh], (unsigned long long)r);
    if (hh) atomicAdd(&sums[2 * sh + 1], (unsigned long long)hh);
  }
}

// Outputs on the 8 QuickCheck vectors only (fp as k_fingerprints writes it),
// eight sequences per pass: lane group g = lane / 8 runs sequence q + g of a
// run of sequences with equal opcodes (per-lane immediates), lane v = lane %
// 8 its vector v.  Runs of one first instruction share their prefix as in
// k_fingerprints.  bytes, cycles, the pruner flag and bad_op as there.
// Cls: the outputs stay in LDS and lane q looks up sequence q's QuickCheck
// class as k_qc_count does (counters into ca.sums, class members appended
// to ca.list with one atomic per block); only class members' outputs are
// written to fp, the only ones read later, so the other 85% of the length-2
// targets cost no fingerprint traffic at all.
struct QcClassArgs {
  const unsigned long long *ckey;
  const uint32_t *cdir;
  uint32_t cdk;
  const uint32_t *crep, *ccnt;
  const uint2 *cfp;
  uint32_t lo_mask;
  unsigned long long *sums;
  uint32_t *list, *list_n;
  uint32_t *tcls;  // the class of each class member (others are not written)
};
// A wave takes fpq_per_wave(len) sequences (their words one per lane).
__host__ __device__ constexpr uint32_t fpq_per_wave(uint32_t len) { return len <= 2 ? 32u : 16u; }
template <bool Cls>
__global__ __launch_bounds__(256) void k_fp_qc(const uint32_t *__restrict__ words, uint32_t len, uint64_t n,
                                               int bytes_model, uint2 *__restrict__ fp, uint8_t *bytes, uint8_t *cyc,
                                               uint8_t *prune, uint32_t *bad_op, QcClassArgs ca) {
  __shared__ uint2 s_fp[Cls ? 4 : 1][Cls ? 32 : 1][kFpVecs];
  __shared__ unsigned long long s_raw[4], s_hit[4];
  __shared__ uint32_t s_in[4], s_base;
  const uint32_t per = fpq_per_wave(len);  // len <= 4
  const uint32_t wid = uni(threadIdx.x >> 6);
  const uint64_t first = ((uint64_t)blockIdx.x * 4 + wid) * per;  // wave-uniform
  if (!Cls && first >= n) return;
  const uint32_t lane = lane_id(), v = lane & 7, g = lane >> 3;
  const uint32_t cnt = first >= n ? 0u : (uint32_t)(first + per < n ? per : n - first);
  const uint32_t wv = lane < cnt * len ? words[first * len + lane] : 0u;
  const uint32_t cmask = cnt >= 32 ? 0xFFFFFFFFu : (1u << cnt) - 1u;
  uint32_t bad_seqs, same_next;
  bool share;
  {
    const uint32_t q = lane < cnt ? lane : 0u, qn = q + 1 < cnt ? q + 1 : q;
    bool l_bad = false, l_same = q + 1 < cnt, l_pre = true;
    for (uint32_t k = 0; k < len; k++) {
      const uint32_t w = (uint32_t)__shfl((int)wv, (int)(q * len + k));
      const uint32_t wn = (uint32_t)__shfl((int)wv, (int)(qn * len + k));
      l_bad |= (w & 0xFFFFu) >= OP_COUNT;
      l_same &= (w & 0xFFFFu) == (wn & 0xFFFFu);
      if (k + 1 < len) l_pre &= w == __builtin_amdgcn_readlane(wv, k);
    }
    bad_seqs = (uint32_t)ballot(l_bad) & cmask;
    same_next = (uint32_t)ballot(l_same) & cmask;
    share = len > 1 && ((uint32_t)ballot(l_pre) & cmask) == cmask;
  }
  State base = load_vec(v);
  uint32_t k0 = 0;
  if (share) {
    for (uint32_t k = 0; k + 1 < len; k++) {
      const uint32_t w = __builtin_amdgcn_readlane(wv, k);
      step(base, w & 0xFFFFu, w >> 16);
    }
    k0 = len - 1;
  }
  for (uint32_t q = 0; q < cnt;) {
    // sequences q .. q + r - 1 have the opcodes of q (so all or none is bad)
    uint32_t r = (uint32_t)__builtin_ctz(~(same_next >> q)) + 1u;
    r = min(min(r, 8u), cnt - q);
    if (!((bad_seqs >> q) & 1u)) {
      const uint32_t sq = q + min(g, r - 1u);
      State st = base;
      for (uint32_t k = k0; k < len; k++) {
        const uint32_t op = __builtin_amdgcn_readlane(wv, q * len + k) & 0xFFFFu;
        const uint32_t w = (uint32_t)__shfl((int)wv, (int)(sq * len + k));
        step(st, op, w >> 16);
      }
      st.p &= kPMask;
      if (g < r) {
        const uint2 o = make_uint2(pack_lo(st), pack_hi(st));
        if constexpr (Cls) s_fp[wid][q + g][v] = o;
        else fp[(first + q + g) * kFpVecs + v] = o;
      }
    }
    q += r;
  }
  if (bad_seqs && lane == 0) atomicOr(bad_op, 1u);
  const uint64_t seq = first + lane;
  bool good = false;  // lane q: sequence q has known opcodes
  uint32_t my_bytes = 0;
  if (lane < cnt) {
    uint32_t w[4] = {0, 0, 0, 0};
    bool bad = false;
    for (uint32_t k = 0; k < len; k++) {
      w[k] = words[seq * len + k];
      bad |= (w[k] & 0xFFFFu) >= OP_COUNT;
    }
    if (!bad) {
      int b = 0, c = 0;
      for (uint32_t k = 0; k < len; k++) {
        b += byte_size(w[k] & 0xFFFFu, bytes_model);
        c += cycles(w[k] & 0xFFFFu);
      }
      bytes[seq] = (uint8_t)b;
      cyc[seq] = (uint8_t)c;
      if (prune) prune[seq] = should_prune(w, (int)len, c_reads, c_writes) ? 1 : 0;
      good = true;
      my_bytes = (uint32_t)b;
    }
  }
  if constexpr (Cls) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    uint64_t raw = 0, hit = 0;
    bool in_class = false;
    uint2 o[8];
    if (good) {
#pragma unroll
      for (int k = 0; k < 8; k++) o[k] = s_fp[wid][lane][k];
      uint32_t k1, k2;
      qc_keys(o, ca.lo_mask, k1, k2);
      const unsigned long long key = (unsigned long long)k1 | ((unsigned long long)k2 << 32);
      const uint32_t b = (uint32_t)(key >> (64 - ca.cdk));
      uint32_t lo = ca.cdir[b];
      const uint32_t hi = ca.cdir[b + 1];
      while (lo < hi && ca.ckey[lo] < key) lo++;
      if (lo < hi && ca.ckey[lo] == key) {
        uint2 r[8];
        load_fp8(ca.cfp, ca.crep[lo], r);
        bool okm, okf;
        fp_compare(o, r, ca.lo_mask, okm, okf);
        in_class = okm;
        if (okm) {
          const uint32_t *c = ca.ccnt + (uint64_t)lo * (kClsBins + 1);
          raw = c[kClsBins];
          const uint32_t tb = min(my_bytes, kClsBins);
          for (uint32_t bb = 0; bb < tb; bb++) hit += c[bb];
          uint4 *dst = reinterpret_cast<uint4 *>(fp + seq * kFpVecs);
#pragma unroll
          for (int k = 0; k < 4; k++) dst[k] = make_uint4(o[2 * k].x, o[2 * k].y, o[2 * k + 1].x, o[2 * k + 1].y);
          ca.tcls[seq] = lo;
        }
      }
    }
    const uint64_t m = ballot(in_class);
    raw = wave_sum64(raw);
    hit = wave_sum64(hit);
    if (lane == 0) {
      s_raw[wid] = raw;
      s_hit[wid] = hit;
      s_in[wid] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint32_t tot = s_in[0] + s_in[1] + s_in[2] + s_in[3];
      s_base = tot ? atomicAdd(ca.list_n, tot) : 0u;
      const uint64_t rr = s_raw[0] + s_raw[1] + s_raw[2] + s_raw[3], hh = s_hit[0] + s_hit[1] + s_hit[2] + s_hit[3];
      const uint32_t sh = blockIdx.x & (kQcCntShards - 1);
      if (rr) atomicAdd(&ca.sums[2 * sh], (unsigned long long)rr);
      if (hh) atomicAdd(&ca.sums[2 * sh + 1], (unsigned long long)hh);
    }
    __syncthreads();
    uint32_t pre = s_base;
    for (uint32_t k = 0; k < wid; k++) pre += s_in[k];
    if (in_class) ca.list[pre + mbcnt(m)] = (uint32_t)seq;
  }
}

// The join: a wave takes 64 targets, reads each one's bucket of the
// candidates' sorted join keys and walks the prefix-summed items (as
// k_qc_join); an entry with an equal key gets QuickCheck's exact compare and
// MidCheck's test.  Survivors (target, slot | pass << 30) leave through a
// per-wave LDS queue, reserved with one atomic on counters[1] per flush (the
// MidCheck count); entries past out_cap are counted, not written, and the
// caller re-runs the segment with a larger buffer.  cur0 (first hit): the
// first round's minimum, as k_midcheck leaves it.  gridDim.y waves share each
// 64-target chunk, taking its items in turn (blockIdx.y = the wave's turn):
// few targets with many items each (length-3 targets against 8.2M
// candidates) then still fill the device.
__global__ __launch_bounds__(256) void k_qcm_join(const uint2 *__restrict__ tfp, const uint2 *__restrict__ tmh,
                                                  const uint8_t *__restrict__ tbytes, uint32_t nt, uint32_t lo_mask,
                                                  const uint32_t *__restrict__ sj, const uint32_t *__restrict__ sjslot,
                                                  const uint32_t *__restrict__ dir, uint32_t dk,
                                                  const uint2 *__restrict__ cfp, const uint2 *__restrict__ cmh,
                                                  const uint8_t *__restrict__ cbytes,
                                                  const uint8_t *__restrict__ cprune, int hits_all,
                                                  uint2 *__restrict__ out, uint32_t out_cap,
                                                  uint32_t *__restrict__ mcnt, const uint32_t *__restrict__ rank,
                                                  uint32_t *cur0, const uint32_t *__restrict__ list,
                                                  const uint32_t *__restrict__ list_n, const uint32_t *__restrict__ tcls,
                                                  const uint32_t *__restrict__ ccls, const uint2 *__restrict__ cpsig) {
  __shared__ uint2 queue[4][kQcQueue];
  __shared__ uint32_t s_pre[4][kWave], s_lo[4][kWave], s_key[4][kWave], s_tid[4][kWave];
  __shared__ uint32_t s_cls[4][kWave];
  __shared__ uint2 s_ps[4][kWave];
  const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
  uint2 *q = queue[w];
  uint32_t qn = 0;
  // this wave's survivor region: counter mcnt[region * kMidCntStride], entries from region * out_cap
  const uint32_t region = ((blockIdx.y * gridDim.x + blockIdx.x) * 4 + w) % kMidRegions;
  uint2 *rout = out + (uint64_t)region * out_cap;
  auto flush = [&]() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(&mcnt[region * kMidCntStride], qn);
    base = uni(base);
    for (uint32_t i = lane; i < qn; i += kWave)
      if (base + i < out_cap) rout[base + i] = q[i];
    __builtin_amdgcn_wave_barrier();
    qn = 0;
  };
  // without a list: targets 0 .. nt, 64 per wave; with one: list[0 ..
  // *list_n), 64 entries per wave, the waves striding over the list
  const uint32_t n = list ? uni(*list_n) : nt;
  const uint32_t stride = list ? gridDim.x * 4 * kWave : 0u;
  for (uint32_t i0 = (blockIdx.x * 4 + w) * kWave; i0 < n; i0 += stride) {
  const bool have = i0 + lane < n;
  const uint32_t t = !have ? 0u : list ? list[i0 + lane] : i0 + lane;
  uint32_t jk = 0, cls = kNoClass;
  uint2 ps = make_uint2(0, 0);
  if (have) {
    uint2 o[8];
    load_fp8(tfp, t, o);
    uint32_t k1, k2;
    qc_keys(o, lo_mask, k1, k2);
    jk = join_key(k1, tmh[t].y);
    cls = tcls[t];
    ps = p_sig(o);
  }
  const uint32_t b = jk >> (32 - dk);
  const uint32_t lo = have ? dir[b] : 0u, c = have ? dir[b + 1] - lo : 0u;
  const uint32_t pre = wave_excl_scan(c, lane);
  const uint32_t total = __builtin_amdgcn_readlane(pre + c, kWave - 1);
  __builtin_amdgcn_wave_barrier();
  s_pre[w][lane] = pre;
  s_lo[w][lane] = lo;
  s_key[w][lane] = jk;
  s_tid[w][lane] = t;
  s_cls[w][lane] = cls;
  s_ps[w][lane] = ps;
  __builtin_amdgcn_wave_barrier();
  for (uint32_t base = blockIdx.y * kWave; base < total; base += kWave * gridDim.y) {
    const uint32_t i = base + lane;
    bool ok = false;
    uint32_t tt = 0, slot = 0, pass = 1;
    if (i < total) {
      uint32_t o = 0;
#pragma unroll
      for (uint32_t step = kWave / 2; step >= 1; step >>= 1)
        if (s_pre[w][o + step] <= i) o += step;
      const uint32_t e = s_lo[w][o] + (i - s_pre[w][o]);
      if (sj[e] == s_key[w][o]) {
        tt = s_tid[w][o];
        slot = sjslot[e];
        // QuickCheck by class: the masked compare holds iff target and
        // candidate are in one class (every member of a class equals its
        // first under the mask); then the exact compare holds iff their P
        // bytes agree too
        const bool okm = ccls[slot] == s_cls[w][o];
        const uint2 cps = cpsig[slot], tps = s_ps[w][o];
        const bool okf = okm && cps.x == tps.x && cps.y == tps.y;
        const bool qc_hit = okm && cbytes[slot] < tbytes[tt] && cprune[slot] == 0;
        const bool exact = !hits_all && okf;  // pass 1 compares P under 0xCF
        pass = (hits_all || okf) ? 1u : 2u;
        const uint2 ma = tmh[tt], mb = cmh[slot];
        ok = qc_hit && (exact ? ma.x == mb.x : ma.y == mb.y);
      }
    }
    const uint64_t m = ballot(ok);
    if (m) {
      if (ok) q[qn + mbcnt(m)] = make_uint2(tt, slot | (pass << 30));
      qn += (uint32_t)__popcll(m);
      if (qn > kQcQueue - kWave) flush();
    }
    if (cur0) wave_run_min(ok ? tt * 2 + (pass - 1) : kRankNone, ok ? rank[slot] : 0u, cur0);
  }
  if (!list) break;
  }
  if (qn) flush();
}

// The survivor regions' counters (kMidCntStride apart) gathered into one
// block, so that the host reads them in one short copy.
__global__ void k_gather_counts(const uint32_t *mcnt, uint32_t *out) {
  if (threadIdx.x < kMidRegions) out[threadIdx.x] = mcnt[threadIdx.x * kMidCntStride];
}

// A segment's first-hit state, cleared in one launch: cur = none, rej = 0,
// done = 0 and, when given, the rule slots' flags = 0.
__global__ void k_seg_clear(uint32_t *cur, uint32_t *rej, uint8_t *done, uint32_t *rule_flag, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  cur[i] = kRankNone;
  rej[i] = 0;
  done[i] = 0;
  if (rule_flag) rule_flag[i] = 0;
}

// First-hit rounds.  cur/rej/done are per (target, pass-1): cur = lowest rank
// this round, rej = 1 + highest rank rejected so far, done = resolved.
__global__ __launch_bounds__(256) void k_round_min(const uint2 *mid, MidList ml,
                                                   const uint32_t *rank, uint32_t *cur,
                                                   const uint32_t *rej, const uint8_t *done) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  uint32_t key = kRankNone, r = 0;
  if (i < ml.size()) {
    const uint2 e = mid[ml.at(i)];
    const uint32_t slot = e.y & 0x3FFFFFFFu, k = e.x * 2 + (e.y >> 30) - 1;
    r = rank[slot];
    if (!done[k] && r + 1 > rej[k]) key = k;
  }
  wave_run_min(key, r, cur);
}

constexpr int kJobWords = 10;  // longest sequence a job holds (STOKE candidates)

struct Job {
  uint32_t tw[kJobWords];
  uint32_t cw[kJobWords];
  uint8_t tl, cl, pass, pmask;
  uint32_t target;  // index within the segment (or pair index for sx_check_pairs)
  uint32_t slot;
  uint8_t nsets, pad[3];
  uint8_t sets[8];
  uint8_t logsz[8];
};

__device__ __forceinline__ void fill_job(Job &j, const uint32_t *targets, uint32_t tl,
                                         const uint32_t *cwords, uint32_t cl, uint32_t t,
                                         uint32_t slot, uint32_t pass, uint32_t pmask) {
  for (int k = 0; k < kJobWords; k++) {
    j.tw[k] = k < (int)tl ? targets[(uint64_t)t * tl + k] : 0;
    j.cw[k] = 0;
  }
  // candidate words: slot < 2863 are single instructions, the rest pairs
  if (slot < (uint32_t)kNumInsts) {
    j.cw[0] = go_inst(slot);
    j.cl = 1;
  } else {
    uint32_t s2 = slot - kNumInsts;
    j.cw[0] = go_inst(s2 / kNumInsts);
    j.cw[1] = go_inst(s2 % kNumInsts);
    j.cl = 2;
  }
  (void)cwords;
  (void)cl;
  j.tl = (uint8_t)tl;
  j.pass = (uint8_t)pass;
  j.pmask = (uint8_t)pmask;
  j.target = t;
  j.slot = slot;
}

// The sweep plan of a job: its sub-sweeps, and the number of 2^14-state
// chunks returned.  pad[0] marks a job whose sub-sweeps together fit one
// chunk: one wave sweeps them all.
__device__ __forceinline__ uint32_t plan_job(Job &j, int mode) {
  Plan pl = make_plan(j.tw, j.tl, j.cw, j.cl, j.pmask, mode, c_reads, c_dep);
  uint32_t total = 0;
  uint64_t states = 0;
#pragma unroll
  for (int s = 0; s < 8; s++) {  // unused entries are 0
    j.sets[s] = pl.sets[s];
    j.logsz[s] = pl.logsz[s];
    if (s < pl.nsets) {
      total += pl.logsz[s] > kChunkLog ? (1u << (pl.logsz[s] - kChunkLog)) : 1u;
      states += 1ull << pl.logsz[s];
    }
  }
  j.nsets = pl.nsets;
  j.pad[0] = (pl.nsets > 1 && states <= (1ull << kChunkLog)) ? 1 : 0;
  return j.pad[0] ? 1u : total;
}

// Plans of length-2 targets against single-instruction candidates, by
// opcode.  The planner reads only the opcodes (an immediate does not change
// which state bytes an instruction reads or writes), so an enumeration of
// length-2 targets tabulates the 58^3 opcode triples under both compare masks
// (390,224 plans, 9.4 MB) and its job kernels look plans up instead of
// running the planner per job.  Entry i = ((t0 * 58 + t1) * 58 + c) * 2 + m,
// m = 0 for the 0xCF mask and 1 for the relaxed one.
struct PlanEntry {
  uint32_t sets_lo, sets_hi, log_lo, log_hi, chunks, meta;  // meta = nsets | single-chunk << 8
};
constexpr uint32_t kPlanTab = (uint32_t)OP_COUNT * OP_COUNT * OP_COUNT * 2;
__global__ void k_plan_table(int mode, uint32_t pm_masked, PlanEntry *tab) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kPlanTab) return;
  const uint32_t trip = i >> 1, c = trip % OP_COUNT, t1 = (trip / OP_COUNT) % OP_COUNT,
                 t0 = trip / (OP_COUNT * OP_COUNT);
  Job j;
  for (int k = 0; k < kJobWords; k++) {
    j.tw[k] = 0;
    j.cw[k] = 0;
  }
  j.tw[0] = t0;  // immediate 0: the plan does not depend on it
  j.tw[1] = t1;
  j.cw[0] = c;
  j.tl = 2;
  j.cl = 1;
  j.pmask = (uint8_t)((i & 1) ? pm_masked : kPMask);
  PlanEntry e;
  e.chunks = plan_job(j, mode);
  e.sets_lo = j.sets[0] | j.sets[1] << 8 | j.sets[2] << 16 | (uint32_t)j.sets[3] << 24;
  e.sets_hi = j.sets[4] | j.sets[5] << 8 | j.sets[6] << 16 | (uint32_t)j.sets[7] << 24;
  e.log_lo = j.logsz[0] | j.logsz[1] << 8 | j.logsz[2] << 16 | (uint32_t)j.logsz[3] << 24;
  e.log_hi = j.logsz[4] | j.logsz[5] << 8 | j.logsz[6] << 16 | (uint32_t)j.logsz[7] << 24;
  e.meta = j.nsets | (uint32_t)j.pad[0] << 8;
  tab[i] = e;
}
// plan_job through the table when it covers the job (tab given, a length-2
// target, a single-instruction candidate, opcodes in range, one of the two
// masks); otherwise the planner.
__device__ __forceinline__ uint32_t plan_job_tab(Job &j, int mode, const PlanEntry *tab, uint32_t pm_masked) {
  if (tab && j.tl == 2 && j.cl == 1) {
    const uint32_t t0 = j.tw[0] & 0xFFFFu, t1 = j.tw[1] & 0xFFFFu, c = j.cw[0] & 0xFFFFu;
    const bool m0 = j.pmask == kPMask, m1 = j.pmask == pm_masked;
    if (t0 < (uint32_t)OP_COUNT && t1 < (uint32_t)OP_COUNT && c < (uint32_t)OP_COUNT && (m0 || m1)) {
      const PlanEntry e = tab[((t0 * OP_COUNT + t1) * OP_COUNT + c) * 2 + (m0 ? 0u : 1u)];
#pragma unroll
      for (int s = 0; s < 4; s++) {
        j.sets[s] = (uint8_t)(e.sets_lo >> (8 * s));
        j.sets[s + 4] = (uint8_t)(e.sets_hi >> (8 * s));
        j.logsz[s] = (uint8_t)(e.log_lo >> (8 * s));
        j.logsz[s + 4] = (uint8_t)(e.log_hi >> (8 * s));
      }
      j.nsets = (uint8_t)e.meta;
      j.pad[0] = (uint8_t)(e.meta >> 8);
      return e.chunks;
    }
  }
  return plan_job(j, mode);
}

constexpr uint32_t kGroupNone = 0xFFFFFFFFu;
constexpr uint32_t kGroupKeyBits = 19;  // kPlanTab < 2^19; kGroupNone sorts last
static_assert(kPlanTab < (1u << kGroupKeyBits), "group keys");
constexpr uint32_t kGroupSort = (1u << kGroupKeyBits) - 1u;  // the sort key of ineligible jobs
// Keyed jobs touch only A, X, Y and P (their opcodes and their sub-sweeps).
constexpr uint32_t kRegsAXYP = (1u << RA) | (1u << RX) | (1u << RY) | (1u << RP);

// The group key of a planned job from its words (Job layout: tw 0-9, cw
// 10-19, word 20 = tl | cl << 8 | pass << 16 | pmask << 24, 23 = nsets, 24-25
// the sets, 26-27 their log sizes).
__device__ __forceinline__ uint32_t group_key_words(uint32_t tw0, uint32_t tw1, uint32_t cw0, uint32_t hdr,
                                                    uint32_t nsets, uint64_t sets8, uint64_t logs8) {
  const uint32_t tl = hdr & 0xFFu, cl = (hdr >> 8) & 0xFFu, pm = hdr >> 24;
  const uint32_t t0 = tw0 & 0xFFFFu, t1 = tw1 & 0xFFFFu, c = cw0 & 0xFFFFu;
  bool ok = tl == 2 && cl == 1 && t0 < (uint32_t)OP_COUNT && t1 < (uint32_t)OP_COUNT && c < (uint32_t)OP_COUNT &&
            nsets >= 1 && nsets <= 8;
  uint32_t touch = ok ? pk_touch(t0) | pk_touch(t1) | pk_touch(c) : 0u;
  for (uint32_t s = 0; ok && s < nsets; s++) {
    ok = ((logs8 >> (8 * s)) & 0xFFu) <= 9u;
    touch |= (uint32_t)(sets8 >> (8 * s)) & 0xFFu;
  }
  ok = ok && (touch & ~kRegsAXYP) == 0;
  return ok ? ((t0 * OP_COUNT + t1) * OP_COUNT + c) * 2u + (pm != kPMask ? 1u : 0u) : kGroupSort;
}
__device__ __forceinline__ uint32_t group_key(const Job &j) {
  uint64_t sets8, logs8;
  __builtin_memcpy(&sets8, j.sets, 8);
  __builtin_memcpy(&logs8, j.logsz, 8);
  const uint32_t hdr = (uint32_t)j.tl | (uint32_t)j.cl << 8 | (uint32_t)j.pass << 16 | (uint32_t)j.pmask << 24;
  return group_key_words(j.tw[0], j.tw[1], j.cw[0], hdr, j.nsets, sets8, logs8);
}
// A planned job's group key into key/idx (when the grouped sweep is on): a
// keyed job leaves the state-parallel sweep (zero chunks).
__device__ __forceinline__ void write_group_key(const Job &j, uint32_t i, uint32_t *chunks, uint32_t *key,
                                                uint32_t *idx) {
  const uint32_t k = group_key(j);
  key[i] = k;
  idx[i] = i;
  if (k != kGroupSort) chunks[i] = 0;
}

// Jobs of one round: the current candidate of every open (target, pass).
// In first-hit mode it also counts the survivors not yet tested (*nrest),
// which decides when the tail runs as one batch.  plan: plan the jobs here
// (all hits, where every lane takes a job); in first-hit mode one survivor in
// several takes one, so the planner would run on mostly idle waves and the
// jobs are planned afterwards, densely, by k_exh_plan.
__global__ __launch_bounds__(kBlk) void k_round_jobs(const uint2 *mid, MidList ml, const uint32_t *rank,
                                                     const uint32_t *cur, const uint32_t *rej, const uint8_t *done,
                                                     const uint32_t *targets, uint32_t tl, uint32_t pm_masked,
                                                     int hits_all, Job *jobs, uint32_t job_cap, uint32_t *njobs,
                                                     uint32_t *nrest, int plan, int plan_mode,
                                                     const PlanEntry *ptab, uint32_t *chunks, uint32_t *fail,
                                                     uint32_t *gkey, uint32_t *gidx) {
  __shared__ uint32_t s_rest[kBlkWaves], s_take[kBlkWaves];
  __shared__ uint32_t s_base;
  const uint32_t i = blockIdx.x * kBlk + threadIdx.x;
  bool take = false, rest = false;
  uint2 e = make_uint2(0, 0);
  if (i < ml.size()) {
    e = mid[ml.at(i)];
    if (hits_all) {
      take = true;
    } else {
      uint32_t slot = e.y & 0x3FFFFFFFu, p = (e.y >> 30) - 1;
      uint64_t k = (uint64_t)e.x * 2 + p;
      take = cur[k] == rank[slot];
      rest = !done[k] && rank[slot] >= rej[k];
    }
  }
  const uint64_t r = __ballot(rest), m = __ballot(take);
  uint32_t tr = 0, tt = 0;
  block_prefix((uint32_t)__popcll(r), s_rest, &tr);
  const uint32_t pre = block_prefix((uint32_t)__popcll(m), s_take, &tt);
  if (threadIdx.x == 0) {
    if (tr) atomicAdd(nrest, tr);
    if (tt) s_base = atomicAdd(njobs, tt);
  }
  __syncthreads();
  if (take) {
    uint32_t slot = e.y & 0x3FFFFFFFu, pass = e.y >> 30;
    uint32_t pm = (hits_all || pass == 2) ? pm_masked : kPMask;
    Job j;
    fill_job(j, targets, tl, nullptr, 0, e.x, slot, pass, pm);
    uint32_t o = s_base + pre + mbcnt(m);
    if (o < job_cap) {
      if (plan) {
        chunks[o] = plan_job_tab(j, plan_mode, ptab, pm_masked);
        fail[o] = 0;
        if (gkey) write_group_key(j, o, chunks, gkey, gidx);
      }
      jobs[o] = j;
    }
  }
}

// The first round's jobs (first hit), one lane per (target, pass) entry k
// whose minimum the MidCheck kernel recorded (cur[k] != none) instead of one
// per survivor; rank == slot (candidate slots are numbered in Go order), so
// the minimum names the candidate.  The jobs are planned afterwards, densely,
// by k_exh_plan (planning here, even by the table, made this kernel slower
// than the two together: most of its lanes take no job).
__global__ __launch_bounds__(kBlk) void k_entry_jobs(const uint32_t *cur, uint32_t nk, const uint32_t *targets,
                                                     uint32_t tl, uint32_t pm_masked, Job *jobs, uint32_t job_cap,
                                                     uint32_t *njobs) {
  __shared__ uint32_t s_take[kBlkWaves];
  __shared__ uint32_t s_base;
  const uint32_t k = blockIdx.x * kBlk + threadIdx.x;
  const uint32_t slot = k < nk ? cur[k] : kRankNone;
  const bool take = slot != kRankNone;
  const uint64_t m = __ballot(take);
  uint32_t tt = 0;
  const uint32_t pre = block_prefix((uint32_t)__popcll(m), s_take, &tt);
  if (threadIdx.x == 0 && tt) s_base = atomicAdd(njobs, tt);
  __syncthreads();
  if (take) {
    const uint32_t pass = (k & 1) + 1;
    Job j;
    fill_job(j, targets, tl, nullptr, 0, k >> 1, slot, pass, pass == 2 ? pm_masked : kPMask);
    const uint32_t o = s_base + pre + mbcnt(m);
    if (o < job_cap) jobs[o] = j;
  }
}

// Plan each job's sweep and count its chunks; with gkey/gidx, also the
// job's group key for the grouped sweep (k_group_keys' work, while the job is
// in registers).
__global__ __launch_bounds__(256) void k_exh_plan(Job *jobs, uint32_t nj, int mode, const PlanEntry *ptab,
                                                  uint32_t pm_masked, uint32_t *chunks, uint32_t *fail,
                                                  uint32_t *gkey, uint32_t *gidx) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nj) return;
  Job j = jobs[i];
  chunks[i] = plan_job_tab(j, mode, ptab, pm_masked);
  jobs[i] = j;
  fail[i] = 0;
  if (gkey) write_group_key(j, i, chunks, gkey, gidx);
}

// Exhaustive sweep: one wave per chunk of 2^14 states.  The chunks of all
// jobs are numbered consecutively (offs = exclusive scan of the per-job chunk
// counts); a launch covers chunks [chunk0, chunk0 + nchunks).  A mismatch sets
// the job's fail flag, which the job's other waves poll to stop early.
// Chunk -> job map: thread j writes j into the map entries of its chunks,
// with bit 31 set when the job is a single chunk (its wave then needs neither
// the job's chunk offset nor its fail flag).
constexpr uint32_t kMapSingle = 0x80000000u;
__global__ void k_chunk_map(const uint64_t *offs, const uint32_t *chunks, uint32_t nj, uint32_t *map) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nj) return;
  const uint64_t o = offs[j];
  const uint32_t n = chunks[j];
  for (uint32_t k = 0; k < n; k++) map[o + k] = j | (n == 1 ? kMapSingle : 0u);
}
// The same map written per chunk: a binary search over the non-decreasing
// offsets finds the job j with offs[j] <= c < offs[j + 1] (offs[nj] is tot;
// jobs without chunks never satisfy it).  Used for sweeps where a few jobs
// hold most chunks and one thread per job would write thousands of entries.
__global__ void k_chunk_map_search(const uint64_t *offs, const uint32_t *chunks, uint32_t nj, uint64_t tot,
                                   uint32_t *map) {
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < tot; c += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t lo = 0, hi = nj;  // offs[lo] <= c < offs[hi] (offs[nj] = tot)
    while (hi - lo > 1) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (offs[mid] <= c) lo = mid;
      else hi = mid;
    }
    map[c] = lo | (chunks[lo] == 1 ? kMapSingle : 0u);
  }
}

// A job as 28 dwords, one per lane of a wave (lane l holds dword l): words
// by readlane instead of dependent scalar loads.
constexpr int kJobDwords = (int)(sizeof(Job) / 4);
static_assert(sizeof(Job) == 112 && kJobDwords <= kWave, "job layout");
static_assert(offsetof(Job, tw) == 0 && offsetof(Job, cw) == 40 && offsetof(Job, tl) == 80 &&
                  offsetof(Job, nsets) == 92 && offsetof(Job, sets) == 96 && offsetof(Job, logsz) == 104,
              "job layout");
__device__ __forceinline__ uint32_t job_byte(uint32_t jv, uint32_t dword0, uint32_t i) {
  return (__builtin_amdgcn_readlane(jv, dword0 + (i >> 2)) >> (8 * (i & 3u))) & 0xFFu;
}

// Sub-sweep fields: varying register r occupies wd[r] bits (8, 1 for the
// carry, 0 when fixed) at bit sh[r] of the state index.
struct ExhFields {
  uint32_t sh[8], wd[8];
};

// One pass of a sub-sweep: lane L evaluates states it + j * 64 + L (j < K)
// below count, each opcode of both sequences dispatched once for the K
// states.  jv is the job spread over the lanes.  Returns true (and raises
// the job's fail flag) on a mismatch.  Wide: the index has more than 32 bits.
template <int K, bool Wide>
__device__ __forceinline__ bool exh_pass(uint32_t jv, uint32_t tl, uint32_t cl, uint32_t pm, const ExhFields &f,
                                         uint64_t start, uint32_t it, uint32_t count, uint32_t lane, uint32_t ji,
                                         uint32_t *fail) {
  State ts[K], cs[K];
  bool act[K];
#pragma unroll
  for (int j = 0; j < K; j++) {
    const uint32_t k = it + j * kWave + lane;
    act[j] = k < count;
    const uint32_t kk = act[j] ? k : 0u;
    State s;
    if (Wide) {
      const uint64_t idx = start + kk;
      s.a = f.wd[RA] ? (uint32_t)(idx >> f.sh[RA]) & 0xFFu : 0u;
      s.x = f.wd[RX] ? (uint32_t)(idx >> f.sh[RX]) & 0xFFu : 0u;
      s.y = f.wd[RY] ? (uint32_t)(idx >> f.sh[RY]) & 0xFFu : 0u;
      s.p = f.wd[RP] ? (uint32_t)(idx >> f.sh[RP]) & 0x1u : 0u;
      s.s = f.wd[RS] ? (uint32_t)(idx >> f.sh[RS]) & 0xFFu : 0xFDu;
      s.m = f.wd[RM] ? (uint32_t)(idx >> f.sh[RM]) & 0xFFu : 0u;
      s.s0 = f.wd[RS0] ? (uint32_t)(idx >> f.sh[RS0]) & 0xFFu : 0u;
      s.s1 = f.wd[RS1] ? (uint32_t)(idx >> f.sh[RS1]) & 0xFFu : 0u;
    } else {
      // start < 2^32 is a multiple of the chunk size and kk < the chunk
      // size, so the index is start | kk; a zero width extracts 0
      const uint32_t idx = (uint32_t)start | kk;
      s.a = __builtin_amdgcn_ubfe(idx, f.sh[RA], f.wd[RA]);
      s.x = __builtin_amdgcn_ubfe(idx, f.sh[RX], f.wd[RX]);
      s.y = __builtin_amdgcn_ubfe(idx, f.sh[RY], f.wd[RY]);
      s.p = __builtin_amdgcn_ubfe(idx, f.sh[RP], f.wd[RP]);
      s.s = f.wd[RS] ? __builtin_amdgcn_ubfe(idx, f.sh[RS], 8u) : 0xFDu;
      s.m = __builtin_amdgcn_ubfe(idx, f.sh[RM], f.wd[RM]);
      s.s0 = __builtin_amdgcn_ubfe(idx, f.sh[RS0], f.wd[RS0]);
      s.s1 = __builtin_amdgcn_ubfe(idx, f.sh[RS1], f.wd[RS1]);
    }
    ts[j] = s;
    cs[j] = s;
  }
  for (uint32_t q = 0; q < tl; q++) {
    const uint32_t w = __builtin_amdgcn_readlane(jv, q);
    step_many<K>(ts, w & 0xFFFFu, w >> 16);
  }
  for (uint32_t q = 0; q < cl; q++) {
    const uint32_t w = __builtin_amdgcn_readlane(jv, kJobWords + q);
    step_many<K>(cs, w & 0xFFFFu, w >> 16);
  }
  bool bad = false;
#pragma unroll
  for (int j = 0; j < K; j++) {
    const uint32_t diff = (ts[j].a ^ cs[j].a) | (ts[j].x ^ cs[j].x) | (ts[j].y ^ cs[j].y) |
                          ((ts[j].p ^ cs[j].p) & pm) | (ts[j].s ^ cs[j].s) | (ts[j].m ^ cs[j].m) |
                          (ts[j].s0 ^ cs[j].s0) | (ts[j].s1 ^ cs[j].s1);
    bad |= act[j] && diff != 0;
  }
  if (__ballot(bad)) {
    if (lane == 0) ((volatile uint32_t *)fail)[ji] = 1u;
    return true;
  }
  return false;
}

// One chunk of job ji with the one-state interpreter (jv = the job spread over
// the lanes; single = the job is this one chunk, c its global chunk number).
// Returns the states swept.
template <int K>
__device__ __forceinline__ uint32_t exh_chunk(uint32_t ji, uint32_t jv, bool single, uint64_t c,
                                              const uint64_t *__restrict__ offs, uint32_t *fail, uint32_t lane) {
  uint64_t local = 0;
  if (!single) {
    if (((volatile uint32_t *)fail)[ji]) return 0;
    local = c - offs[ji];
  }
  const uint32_t hdr = __builtin_amdgcn_readlane(jv, 20);  // tl, cl, pass, pmask
  const uint32_t tl = hdr & 0xFFu, cl = (hdr >> 8) & 0xFFu, pm = hdr >> 24;
  const uint32_t h23 = __builtin_amdgcn_readlane(jv, 23);  // nsets, pad[0] = packed
  const uint32_t nsets = h23 & 0xFFu;
  const bool packed = ((h23 >> 8) & 0xFFu) != 0;  // one wave sweeps every sub-sweep of the job
  uint32_t s_first = 0;
  if (!packed) {
    for (uint32_t s = 0; s < nsets; s++) {
      const uint32_t l = job_byte(jv, 26, s);
      const uint64_t nch = l > kChunkLog ? (1ull << (l - kChunkLog)) : 1ull;
      if (local < nch) {
        s_first = s;
        break;
      }
      local -= nch;
    }
  }
  const uint32_t s_last = packed ? nsets : s_first + 1;
  uint32_t done_states = 0;
  bool failed = false;
  for (uint32_t si = s_first; si < s_last && !failed; si++) {
    const uint32_t set = job_byte(jv, 24, si), lg = job_byte(jv, 26, si);
    // bit positions of the varying registers in the state index
    ExhFields f;
    uint32_t pos = 0;
#pragma unroll
    for (int r = 0; r < 8; r++) {
      const uint32_t w = (set >> r) & 1u ? (r == RP ? 1u : 8u) : 0u;
      f.sh[r] = pos;
      f.wd[r] = w;
      pos += w;
    }
    const uint64_t start = packed ? 0 : local << kChunkLog;
    const uint32_t count = lg < kChunkLog ? (1u << lg) : (1u << kChunkLog);
    const uint32_t done_before = done_states;
    if (lg > 32) {
      // index wider than 32 bits: one state per lane and pass
      for (uint32_t it = 0; it < count && !failed; it += kWave) {
        done_states = done_before + min(it + kWave, count);
        failed = exh_pass<1, true>(jv, tl, cl, pm, f, start, it, count, lane, ji, fail);
        if (!failed && (it & (32 * kWave - 1)) == 31 * kWave && ((volatile uint32_t *)fail)[ji]) failed = true;
      }
    } else {
      // the first pass covers 64 states (most failing pairs stop there), the
      // later ones K x 64 with one opcode dispatch per K states
      done_states = done_before + min((uint32_t)kWave, count);
      failed = exh_pass<1, false>(jv, tl, cl, pm, f, start, 0, count, lane, ji, fail);
      for (uint32_t it = kWave; it < count && !failed; it += K * kWave) {
        done_states = done_before + min(it + K * kWave, count);
        failed = exh_pass<K, false>(jv, tl, cl, pm, f, start, it, count, lane, ji, fail);
        // other chunks of this job may have failed: look every 2,048 states
        if (!failed && !single && ((it - kWave) & (32 * kWave - 1)) == 32 * kWave - K * kWave &&
            ((volatile uint32_t *)fail)[ji])
          failed = true;
      }
    }
  }
  return done_states;
}

// Packed sweep of one chunk with every per-wave quantity on the scalar unit.
// jw = the job's 28 words (a wave-uniform address: scalar loads).  For a
// sub-sweep of 2^lg states the varying registers take consecutive bit fields
// of the state index in the order A X Y S M S0 S1 P (the carry last), and the
// index splits into lane bits (0-5), the half bit (6), bit 7, pass bits
// (8-13) and chunk bits (14 and up): disjoint, so a field is the OR of its
// parts.  Each lane builds its fields from its lane, half and chunk bits with
// bit 7 clear and set once per sub-sweep (base[0], base[1]).  The first field
// is 8 bits wide whenever the sweep has more than 256 states, so the pass bits
// land in one field only, the second (R2, from bit 8): a pass ORs (it >> 8)
// into R2 and leaves the others at their base.  Fields lie in bits [0, lg), so
// for lg < 7 the lanes past 2^lg repeat states of lanes below it (no activity
// mask is needed) and for lg >= 7 every pass is full.  KP packed pairs per lane
// and opcode dispatch: KP = 1 takes base[(it >> 7) & 1] per pass of 128 states,
// KP = 2 both bases per pass of 256.  (The instruction words by scalar loads
// or by readlane from a register measured the same.)
template <int KP>
__device__ __forceinline__ uint32_t exh_chunk_pk(const uint32_t *__restrict__ jw, uint32_t ji, bool single,
                                                 uint64_t c, const uint64_t *__restrict__ offs, uint32_t *fail,
                                                 uint32_t lane) {
  static_assert(KP == 1 || KP == 2, "pairs per lane");
  uint64_t local = 0;
  if (!single) {
    if (uni(((volatile uint32_t *)fail)[ji])) return 0;
    local = uni64(c - offs[ji]);
  }
  const uint32_t hdr = jw[20];  // tl, cl, pass, pmask
  const uint32_t tl = hdr & 0xFFu, cl = (hdr >> 8) & 0xFFu, pm = hdr >> 24;
  const uint32_t h23 = jw[23];  // nsets, pad[0] = packed
  const uint32_t nsets = h23 & 0xFFu;
  const bool packed = ((h23 >> 8) & 0xFFu) != 0;  // one wave sweeps every sub-sweep of the job
  const uint64_t sets8 = (uint64_t)jw[24] | ((uint64_t)jw[25] << 32);
  const uint64_t logs8 = (uint64_t)jw[26] | ((uint64_t)jw[27] << 32);
  uint32_t s_first = 0;
  if (!packed) {
    for (uint32_t s = 0; s < nsets; s++) {
      const uint32_t l = (uint32_t)(logs8 >> (8 * s)) & 0xFFu;
      const uint64_t nch = l > kChunkLog ? (1ull << (l - kChunkLog)) : 1ull;
      if (local < nch) {
        s_first = s;
        break;
      }
      local -= nch;
    }
  }
  const uint32_t s_last = packed ? nsets : s_first + 1;
  const pk16 pm2 = pk_splat(pm);
  uint32_t done_states = 0;
  for (uint32_t si = s_first; si < s_last; si++) {
    const uint32_t set = (uint32_t)(sets8 >> (8 * si)) & 0xFFu, lg = (uint32_t)(logs8 >> (8 * si)) & 0xFFu;
    const uint32_t count = lg < kChunkLog ? (1u << lg) : (1u << kChunkLog);
    if (lg > 32) {
      // index wider than 32 bits: one state per lane and pass (register order)
      ExhFields f;
      uint32_t pos = 0;
#pragma unroll
      for (int r = 0; r < 8; r++) {
        const uint32_t w = (set >> r) & 1u ? (r == RP ? 1u : 8u) : 0u;
        f.sh[r] = pos;
        f.wd[r] = w;
        pos += w;
      }
      const uint64_t start = local << kChunkLog;
      const uint32_t jw_l = lane < (uint32_t)kJobDwords ? jw[lane] : 0u;
      for (uint32_t it = 0; it < count; it += kWave) {
        done_states += min((uint32_t)kWave, count - it);
        if (exh_pass<1, true>(jw_l, tl, cl, pm, f, start, it, count, lane, ji, fail)) return done_states;
        if ((it & (32 * kWave - 1)) == 31 * kWave && uni(((volatile uint32_t *)fail)[ji])) return done_states;
      }
      continue;
    }
    // field positions in the order A X Y S M S0 S1 P; sel[r] = all ones for R2
    uint32_t sh[8], wd[8], sel[8];
    {
      constexpr int kOrder[8] = {RA, RX, RY, RS, RM, RS0, RS1, RP};
      uint32_t pos = 0;
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const int r = kOrder[k];
        const uint32_t w = (set >> r) & 1u ? (r == RP ? 1u : 8u) : 0u;
        sh[r] = pos;
        wd[r] = w;
        sel[r] = (w && pos == 8u) ? 0xFFFFFFFFu : 0u;
        pos += w;
      }
    }
    const uint32_t st32 = packed ? 0u : (uint32_t)(local << kChunkLog);
    PState base[2];
#pragma unroll
    for (int b = 0; b < 2; b++) {
      const uint32_t lo = st32 | lane | ((uint32_t)b << 7), hi = lo | (uint32_t)kWave;
      auto fld = [&](int r) {
        return __builtin_bit_cast(pk16, ((lo >> sh[r]) & ((1u << wd[r]) - 1u)) |
                                            (((hi >> sh[r]) & ((1u << wd[r]) - 1u)) << 16));
      };
      base[b].a = fld(RA);
      base[b].x = fld(RX);
      base[b].y = fld(RY);
      base[b].p = fld(RP);
      base[b].s = wd[RS] ? fld(RS) : pk_splat(0xFD);
      base[b].m = fld(RM);
      base[b].s0 = fld(RS0);
      base[b].s1 = fld(RS1);
    }
    for (uint32_t it = 0; it < count; it += 2 * KP * kWave) {
      // the pass bits' value in R2, in both halves, on the scalar unit
      const uint32_t inc = (it >> 8) * 0x10001u;
      PState ts[KP], cs[KP];
      auto init = [&](const PState &bs) {
        PState s;
        s.a = bs.a | __builtin_bit_cast(pk16, inc & sel[RA]);
        s.x = bs.x | __builtin_bit_cast(pk16, inc & sel[RX]);
        s.y = bs.y | __builtin_bit_cast(pk16, inc & sel[RY]);
        s.p = bs.p | __builtin_bit_cast(pk16, inc & sel[RP]);
        s.s = bs.s | __builtin_bit_cast(pk16, inc & sel[RS]);
        s.m = bs.m | __builtin_bit_cast(pk16, inc & sel[RM]);
        s.s0 = bs.s0 | __builtin_bit_cast(pk16, inc & sel[RS0]);
        s.s1 = bs.s1 | __builtin_bit_cast(pk16, inc & sel[RS1]);
        return s;
      };
      if (KP == 2) {
        ts[0] = init(base[0]);
        ts[KP - 1] = init(base[1]);
      } else if ((it >> 7) & 1u) {
        ts[0] = init(base[1]);
      } else {
        ts[0] = init(base[0]);
      }
#pragma unroll
      for (int j = 0; j < KP; j++) cs[j] = ts[j];
      for (uint32_t q = 0; q < tl; q++) {
        const uint32_t w = jw[q];
        step_pk<KP>(ts, w & 0xFFFFu, w >> 16);
      }
      for (uint32_t q = 0; q < cl; q++) {
        const uint32_t w = jw[kJobWords + q];
        step_pk<KP>(cs, w & 0xFFFFu, w >> 16);
      }
      bool bad = false;
#pragma unroll
      for (int j = 0; j < KP; j++) {
        const pk16 d = (ts[j].a ^ cs[j].a) | (ts[j].x ^ cs[j].x) | (ts[j].y ^ cs[j].y) |
                       ((ts[j].p ^ cs[j].p) & pm2) | (ts[j].s ^ cs[j].s) | (ts[j].m ^ cs[j].m) |
                       (ts[j].s0 ^ cs[j].s0) | (ts[j].s1 ^ cs[j].s1);
        bad |= __builtin_bit_cast(uint32_t, d) != 0;
      }
      done_states += min(2u * KP * kWave, count - it);
      if (__ballot(bad)) {
        if (lane == 0) ((volatile uint32_t *)fail)[ji] = 1u;
        return done_states;
      }
      // other chunks of this job may have failed: look every 2,048 states
      if (!single && (it & (32 * kWave - 1)) == 32 * kWave - 2 * KP * kWave && uni(((volatile uint32_t *)fail)[ji]))
        return done_states;
    }
  }
  return done_states;
}

__device__ __forceinline__ uint32_t job_dword(const Job *jobs, uint32_t ji, uint32_t lane) {
  return lane < (uint32_t)kJobDwords ? reinterpret_cast<const uint32_t *>(jobs + ji)[lane] : 0u;
}

// Exhaustive sweep, state-parallel: one wave per chunk of 2^14 states
// (launching a wave per chunk lets the hardware balance the uneven chunks;
// persistent waves with prefetch and two chunks per wave both measured
// slower).  The chunks of all jobs are numbered consecutively (offs =
// exclusive scan of the per-job chunk counts); a launch covers chunks
// [chunk0, chunk0 + nchunks).  The job is one map load away, or found by a
// 64-ary search over the chunk offsets when there is no map (each step the 64
// lanes test 64 evenly spaced offsets and the range shrinks 64-fold).  Pk:
// the packed interpreter with K pairs per lane and dispatch; else the
// one-state interpreter with K states per lane in the passes after the first.
template <int K, bool Pk>
__global__ __launch_bounds__(256) void k_exh_run(const Job *__restrict__ jobs, uint32_t nj,
                                                 const uint64_t *__restrict__ offs,
                                                 const uint32_t *__restrict__ chunk_job, uint64_t chunk0,
                                                 uint64_t nchunks, uint32_t *fail,
                                                 unsigned long long *evaluated) {
  const uint32_t lane = lane_id();
  const uint64_t w = (uint64_t)blockIdx.x * 4 + uni(threadIdx.x >> 6);
  if (w >= nchunks) return;  // whole wave
  const uint64_t c = chunk0 + w;
  uint32_t lo = 0, hi = nj;
  bool single = false;
  if (chunk_job) {
    const uint32_t e = uni(chunk_job[c]);
    lo = e & ~kMapSingle;
    single = (e & kMapSingle) != 0;
    hi = lo + 1;
  }
  while (hi - lo > 1) {
    const uint32_t step = (hi - lo + kWave - 1) / kWave;
    const uint32_t p = lo + lane * step;
    const uint64_t below = __ballot(p < hi && offs[p < hi ? p : lo] <= c);  // lane 0 always holds
    const uint32_t k = 63u - (uint32_t)__builtin_clzll(below);
    lo = uni(lo + k * step);
    hi = min(hi, lo + step);
  }
  const uint32_t ji = uni(lo);
  uint32_t done;
  if constexpr (Pk)
    done = exh_chunk_pk<K>(reinterpret_cast<const uint32_t *>(jobs + ji), ji, single, c, offs, fail, lane);
  else
    done = exh_chunk<K>(ji, job_dword(jobs, ji, lane), single, c, offs, fail, lane);
  if (lane == 0 && done) atomicAdd(&evaluated[w & (kEvalShards - 1)], (unsigned long long)done);
}

// Sweep of small plans with one job per lane.
//
// Most jobs of length-2 targets against single-instruction candidates have
// sub-sweeps of 256 or 512 states (one 8-bit register, and the carry when it
// matters), so a wave per job spends most of its time on the job's set-up.
// Such jobs are keyed by (opcode triple, compare mask), the plan-table index,
// and sorted by key; wave w takes sorted positions [64 w, 64 w + 64).  Lanes
// whose jobs share a key share the plan and every opcode dispatch (only the
// immediates differ, per lane), so the wave sweeps up to 64 jobs at once:
// every register value of a pass is wave-uniform, and lane L holds its job's
// two states of the pass in the low and high halves.  A wave with lanes of
// several keys (at the ends of the sorted runs) takes them one key after the
// other.  A job fails at its first mismatching state; the wave leaves a key
// once every one of its jobs has failed.  The keys and constants are above
// k_round_jobs, whose planning writes them.

// Key of each job (kGroupSort when it stays with the state-parallel sweep),
// for jobs whose planner did not write it; a keyed job gets zero chunks there.
__global__ __launch_bounds__(256) void k_group_keys(const Job *__restrict__ jobs, uint32_t nj, uint32_t *chunks,
                                                    uint32_t *key, uint32_t *idx) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nj) return;
  const uint32_t *jw = reinterpret_cast<const uint32_t *>(jobs + i);
  const uint32_t k = group_key_words(jw[0], jw[1], jw[kJobWords], jw[20], jw[23] & 0xFFu,
                                     (uint64_t)jw[24] | ((uint64_t)jw[25] << 32),
                                     (uint64_t)jw[26] | ((uint64_t)jw[27] << 32));
  key[i] = k;
  idx[i] = i;
  if (k != kGroupSort) chunks[i] = 0;
}

// The items of one key group (lanes gm of the wave) that wave u of the
// block takes: item it of sub-sweep si when (items before si) + it = u (mod
// 4).  Regs: the state registers the group's opcodes and sub-sweeps touch;
// the others stay at their domain values and the compiler drops them.
// Returns the lanes whose job failed; *evald counts the states swept for
// jobs still open.
template <int K, uint32_t Regs>
__device__ __forceinline__ uint64_t group_items(uint64_t gm, uint32_t u, uint32_t op0, uint32_t op1, uint32_t opc,
                                                pk16 imm0, pk16 imm1, pk16 immc, pk16 pm2, uint32_t nsets,
                                                uint64_t sets8, uint64_t logs8, const uint32_t *fail, uint32_t ji,
                                                uint64_t *evald) {
  uint64_t failed = 0;
  uint32_t item = 0;  // items of the sub-sweeps before si
  for (uint32_t si = 0; si < nsets; si++) {
    const uint32_t set = (uint32_t)(sets8 >> (8 * si)) & 0xFFu, lg = (uint32_t)(logs8 >> (8 * si)) & 0xFFu;
    const uint32_t npairs = lg >= 1 ? 1u << (lg - 1) : 1u;  // two states per lane and pair
    const uint32_t nitems = (npairs + 63u) / 64u;
    for (uint32_t it = (u + 4u - (item & 3u)) & 3u; it < nitems; it += 4) {  // item + it = u (mod 4)
      // the jobs still open (other waves of the block may have failed them)
      const uint64_t live = gm & ~ballot(((volatile const uint32_t *)fail)[ji] != 0) & ~failed;
      if (!live) continue;
      // at most one 8-bit register r1 (index bits 0-7) and the carry (the next bit)
      int r1 = -1;
#pragma unroll
      for (int r = 7; r >= 0; r--)
        if (r != RP && ((set >> r) & 1u)) r1 = r;
      const bool hasp = (set >> RP) & 1u;
      const uint32_t q0 = it * 64u, q1 = min(npairs, q0 + 64u);
      uint64_t gfail = 0;
      for (uint32_t q = q0; q < q1 && (gfail | ~live) != ~0ull; q += K) {
        PState ts[K], cs[K];
#pragma unroll
        for (int j = 0; j < K; j++) {
          const uint32_t qj = q + j < q1 ? q + j : q;  // past the end: a pair already swept
          const uint32_t lo = 2 * qj, hi = lg ? 2 * qj + 1 : 0;
          PState st;
          st.a = st.x = st.y = st.p = st.m = st.s0 = st.s1 = pk_splat(0);
          st.s = pk_splat(0xFD);
          if (hasp) st.p = __builtin_bit_cast(pk16, r1 >= 0 ? (lo >> 8) | ((hi >> 8) << 16) : lo | (hi << 16));
          if (r1 >= 0) {
            const pk16 v = __builtin_bit_cast(pk16, (lo & 0xFFu) | ((hi & 0xFFu) << 16));
            if (r1 == RA) st.a = v;
            if constexpr ((Regs >> RX) & 1u) if (r1 == RX) st.x = v;
            if constexpr ((Regs >> RY) & 1u) if (r1 == RY) st.y = v;
            if constexpr ((Regs >> RS) & 1u) if (r1 == RS) st.s = v;
            if constexpr ((Regs >> RM) & 1u) if (r1 == RM) st.m = v;
            if constexpr ((Regs >> RS0) & 1u) if (r1 == RS0) st.s0 = v;
            if constexpr ((Regs >> RS1) & 1u) if (r1 == RS1) st.s1 = v;
          }
          ts[j] = st;
          cs[j] = st;
        }
        step_pk_imm<K, Regs>(ts, op0, imm0);
        step_pk_imm<K, Regs>(ts, op1, imm1);
        step_pk_imm<K, Regs>(cs, opc, immc);
        bool bad = false;
#pragma unroll
        for (int j = 0; j < K; j++) {
          const pk16 d = (ts[j].a ^ cs[j].a) | (ts[j].x ^ cs[j].x) | (ts[j].y ^ cs[j].y) |
                         ((ts[j].p ^ cs[j].p) & pm2) | (ts[j].s ^ cs[j].s) | (ts[j].m ^ cs[j].m) |
                         (ts[j].s0 ^ cs[j].s0) | (ts[j].s1 ^ cs[j].s1);
          bad |= __builtin_bit_cast(uint32_t, d) != 0;
        }
        *evald += (uint64_t)__builtin_popcountll(live & ~gfail) * min((uint32_t)K, q1 - q) * (lg ? 2u : 1u);
        gfail |= ballot(bad) & live;
      }
      failed |= gfail;
    }
    item += nitems;
  }
  return failed;
}

// A workgroup of 4 waves per 64 sorted positions: each wave takes items u,
// u + 4, ... of the block's plan, an item being up to 128 states (64 pairs)
// of one sub-sweep, so the waves stay short and many; every lane of every
// wave reads the same 64 jobs.  K pairs per lane share one opcode dispatch.
// Keyed groups touch only A, X, Y and P, so the interpreter runs without the
// other registers.
// Ent: the jobs are compact records {target word 0, target word 1, slot |
// pass << 30, target} (jent) instead of Jobs; a keyed job's plan is the plan
// table's entry at its group key, which is the table's index, and its mask
// follows from the pass.
template <int K, bool Ent>
__global__ __launch_bounds__(256) void k_exh_group(const Job *__restrict__ jobs, const uint4 *__restrict__ jent,
                                                   const PlanEntry *__restrict__ ptab, uint32_t pm_masked,
                                                   const uint32_t *__restrict__ skey,
                                                   const uint32_t *__restrict__ sidx, uint32_t n, uint32_t *fail,
                                                   unsigned long long *evaluated) {
  static_assert(K == 1 || K == 2 || K == 4 || K == 8, "pairs per lane");
  const uint32_t lane = lane_id();
  const uint32_t u = uni(threadIdx.x >> 6);
  const uint32_t p = blockIdx.x * kWave + lane;
  const uint32_t key0 = p < n ? skey[p] : kGroupSort;
  const bool mine = key0 != kGroupSort;
  uint64_t todo = ballot(mine);
  if (!todo) return;
  const uint32_t ji = mine ? sidx[p] : 0u;
  uint32_t lw0, lw1, lc0, lpm;
  if constexpr (Ent) {
    const uint4 r = jent[ji];
    lw0 = r.x;
    lw1 = r.y;
    lc0 = go_inst(r.z & 0x3FFFFFFFu);
    lpm = (r.z >> 30) == 2 ? pm_masked : kPMask;
  } else {
    const uint32_t *jl = reinterpret_cast<const uint32_t *>(jobs + ji);
    lw0 = jl[0];
    lw1 = jl[1];
    lc0 = jl[kJobWords];
    lpm = jl[20] >> 24;
  }
  const uint32_t key = key0 | (lpm << 24);  // with the exact compare mask
  const pk16 imm0 = pk_splat(lw0 >> 16), imm1 = pk_splat(lw1 >> 16), immc = pk_splat(lc0 >> 16);
  uint64_t failed = 0, evald = 0;
  while (todo) {
    const uint32_t first = (uint32_t)__builtin_ctzll(todo);
    const uint32_t k = __builtin_amdgcn_readlane(key, first);
    const uint64_t gm = ballot(mine && key == k);
    todo &= ~gm;
    if (!(mine && key == k)) continue;
    // the group's opcodes, mask and plan
    uint32_t op0, op1, opc, pm, nsets;
    uint64_t sets8, logs8;
    if constexpr (Ent) {
      op0 = __builtin_amdgcn_readlane(lw0, first) & 0xFFFFu;
      op1 = __builtin_amdgcn_readlane(lw1, first) & 0xFFFFu;
      opc = __builtin_amdgcn_readlane(lc0, first) & 0xFFFFu;
      pm = k >> 24;
      // (the compiler loads the entry through the vector unit: the words are
      // made uniform so that the sweep's loop control stays scalar)
      const PlanEntry e = ptab[k & 0xFFFFFFu];
      nsets = uni(e.meta) & 0xFFu;
      sets8 = (uint64_t)uni(e.sets_lo) | ((uint64_t)uni(e.sets_hi) << 32);
      logs8 = (uint64_t)uni(e.log_lo) | ((uint64_t)uni(e.log_hi) << 32);
    } else {
      // the first lane's job
      const uint32_t *jw = reinterpret_cast<const uint32_t *>(jobs + __builtin_amdgcn_readlane(ji, first));
      op0 = jw[0] & 0xFFFFu;
      op1 = jw[1] & 0xFFFFu;
      opc = jw[kJobWords] & 0xFFFFu;
      pm = jw[20] >> 24;
      nsets = jw[23] & 0xFFu;
      sets8 = (uint64_t)jw[24] | ((uint64_t)jw[25] << 32);
      logs8 = (uint64_t)jw[26] | ((uint64_t)jw[27] << 32);
    }
    failed |= group_items<K, kRegsAXYP>(gm, u, op0, op1, opc, imm0, imm1, immc, pk_splat(pm), nsets, sets8, logs8,
                                        fail, ji, &evald);
  }
  if ((failed >> lane) & 1u) fail[ji] = 1u;
  // evaluated[0 .. kEvalShards) counts states, the next kEvalShards keyed jobs
  if (lane == 0 && evald)
    atomicAdd(&evaluated[(blockIdx.x * 4 + u) & (kEvalShards - 1)], (unsigned long long)evald);
  const uint64_t keyed = ballot(mine);  // all lanes take part in the ballot
  if (lane == 0 && u == 0)
    atomicAdd(&evaluated[kEvalShards + (blockIdx.x & (kEvalShards - 1))], (unsigned long long)__popcll(keyed));
}

// Resolve a round: a passing job emits its rule and closes (target, pass); a
// failing one records its rank as rejected.  Every job resets cur.
__global__ __launch_bounds__(256) void k_round_resolve(const Job *jobs, uint32_t nj, const uint32_t *fail,
                                                       const uint32_t *rank, uint32_t *cur, uint32_t *rej,
                                                       uint8_t *done, const uint2 *tfp, const uint2 *cfp,
                                                       const uint8_t *tbytes, const uint8_t *tcyc,
                                                       const uint8_t *cbytes, const uint8_t *ccyc,
                                                       int hits_all, uint32_t dead, uint64_t tbase,
                                                       sx_rule *rules, uint32_t rule_cap,
                                                       uint32_t *nrules, uint32_t *rule_flag) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nj) return;
  const Job &j = jobs[i];
  uint64_t k = (uint64_t)j.target * 2 + (j.pass - 1);
  bool ok = !fail[i];
  uint32_t fd = 0;
  if (ok && !hits_all && j.pass == 2) {
    const uint2 *a = tfp + (uint64_t)j.target * kFpVecs, *b = cfp + (uint64_t)j.slot * kFpVecs;
    for (int v = 0; v < 8; v++) fd |= ((a[v].x ^ b[v].x) >> 24) & kPMask;
    ok = fd != 0;  // FlagDiff == 0 is skipped (worker.go:211-213)
  }
  if (!hits_all) {
    cur[k] = kRankNone;
    if (ok) done[k] = 1;
    else rej[k] = rank[j.slot] + 1;
  }
  if (!ok) return;
  // first hit: one rule per (target, pass), written to its own slot k (the
  // slots are compacted in order afterwards); all hits: appended
  uint32_t o;
  if (hits_all) {
    o = atomicAdd(nrules, 1u);
    if (o >= rule_cap) return;
  } else {
    o = (uint32_t)k;
    rule_flag[k] = 1;
  }
  sx_rule r;
  for (int q = 0; q < 4; q++) {
    r.src[q] = j.tw[q];
    r.rep[q] = j.cw[q];
  }
  r.target_index = tbase + j.target;
  r.src_len = j.tl;
  r.rep_len = j.cl;
  r.pass = j.pass;
  r.dead_flags = (uint8_t)(hits_all ? dead : (j.pass == 2 ? fd : 0));
  r.bytes_saved = (int16_t)((int)tbytes[j.target] - (int)cbytes[j.slot]);
  r.cycles_saved = (int16_t)((int)tcyc[j.target] - (int)ccyc[j.slot]);
  rules[o] = r;
}

// First-hit tail in one batch.  After a few rounds most (target, pass)
// entries are resolved; for the rest, every Mid survivor not yet tested
// (rank >= rej) becomes a job at once, and the smallest passing Go rank is
// the first hit (the rounds would reach the same candidate one rank at a
// time).  Every job gets its compact record in jent (target words 0-1, slot
// | pass << 30, target: 16 bytes); with a plan table a full Job (112 bytes)
// is written only for a job the grouped sweep does not take (no group key,
// or no grouped sweep), which the state-parallel sweep reads.
__global__ __launch_bounds__(kBlk) void k_bulk_jobs(const uint2 *mid, MidList ml, const uint32_t *rank,
                                                    const uint32_t *rej, const uint8_t *done,
                                                    const uint32_t *targets, uint32_t tl, uint32_t pm_masked,
                                                    Job *jobs, uint4 *jent, uint32_t job_cap, uint32_t *njobs,
                                                    int plan_mode, const PlanEntry *ptab, uint32_t *chunks,
                                                    uint32_t *fail, uint32_t *gkey, uint32_t *gidx) {
  __shared__ uint32_t s_take[kBlkWaves];
  __shared__ uint32_t s_base;
  const uint32_t i = blockIdx.x * kBlk + threadIdx.x;
  bool take = false;
  uint2 e = make_uint2(0, 0);
  if (i < ml.size()) {
    e = mid[ml.at(i)];
    uint32_t slot = e.y & 0x3FFFFFFFu, p = (e.y >> 30) - 1;
    uint64_t k = (uint64_t)e.x * 2 + p;
    take = !done[k] && rank[slot] >= rej[k];
  }
  const uint64_t m = __ballot(take);
  uint32_t tt = 0;
  const uint32_t pre = block_prefix((uint32_t)__popcll(m), s_take, &tt);
  if (threadIdx.x == 0 && tt) s_base = atomicAdd(njobs, tt);
  __syncthreads();
  if (take) {
    uint32_t slot = e.y & 0x3FFFFFFFu, pass = e.y >> 30;
    Job j;
    fill_job(j, targets, tl, nullptr, 0, e.x, slot, pass, pass == 2 ? pm_masked : kPMask);
    uint32_t o = s_base + pre + mbcnt(m);
    if (o < job_cap) {
      const uint32_t c = plan_job_tab(j, plan_mode, ptab, pm_masked);
      const uint32_t k = gkey ? group_key(j) : kGroupSort;
      if (gkey) {
        gkey[o] = k;
        gidx[o] = o;
      }
      chunks[o] = k != kGroupSort ? 0u : c;  // a keyed job leaves the state-parallel sweep
      fail[o] = 0;
      jent[o] = make_uint4(j.tw[0], j.tw[1], e.y, e.x);
      if (k == kGroupSort || !ptab) jobs[o] = j;  // without the table the grouped sweep reads Jobs
    }
  }
}

// FlagDiff over the 8 QuickCheck vectors from the fingerprints: 0 when a
// non-P byte differs is already excluded by the masked QuickCheck.
__device__ __forceinline__ uint32_t fp_flag_diff(const uint2 *tfp, const uint2 *cfp, uint32_t t, uint32_t slot) {
  const uint2 *a = tfp + (uint64_t)t * kFpVecs, *b = cfp + (uint64_t)slot * kFpVecs;
  uint32_t fd = 0;
  for (int v = 0; v < 8; v++) fd |= ((a[v].x ^ b[v].x) >> 24) & kPMask;
  return fd;
}

// The tail's minimum passing rank per (target, pass), from the jobs' compact
// records.  comp: best holds the complement of the minimum (it starts at 0,
// as rej does at round 0, and takes atomicMax of ~rank); else best starts
// at kRankNone and takes atomicMin.
__global__ __launch_bounds__(256) void k_bulk_min(const uint4 *jent, uint32_t nj, const uint32_t *fail,
                                                  const uint32_t *rank, const uint2 *tfp, const uint2 *cfp,
                                                  uint32_t *best, int comp) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nj || fail[i]) return;
  const uint4 r = jent[i];
  const uint32_t target = r.w, slot = r.z & 0x3FFFFFFFu, pass = r.z >> 30;
  if (pass == 2 && fp_flag_diff(tfp, cfp, target, slot) == 0) return;  // worker.go:211-213
  if (comp)
    atomicMax(&best[(uint64_t)target * 2 + (pass - 1)], ~rank[slot]);
  else
    atomicMin(&best[(uint64_t)target * 2 + (pass - 1)], rank[slot]);
}

// Rules of the (target, pass) entries the tail resolved; rank == slot.
__global__ __launch_bounds__(256) void k_bulk_emit(const uint32_t *best, uint32_t nt, const uint32_t *targets,
                                                   uint32_t tl, const uint2 *tfp, const uint2 *cfp,
                                                   const uint8_t *tbytes, const uint8_t *tcyc,
                                                   const uint8_t *cbytes, const uint8_t *ccyc, uint64_t tbase,
                                                   sx_rule *rules, uint32_t *rule_flag, int comp) {
  uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= 2 * nt) return;
  const uint32_t b = comp ? ~best[k] : best[k];  // comp: see k_bulk_min
  if (b == kRankNone) return;
  const uint32_t t = k >> 1, pass = (k & 1) + 1, slot = b;
  Job j;
  fill_job(j, targets, tl, nullptr, 0, t, slot, pass, 0);
  const uint32_t o = k;
  rule_flag[k] = 1;
  sx_rule r;
  for (int q = 0; q < 4; q++) {
    r.src[q] = j.tw[q];
    r.rep[q] = j.cw[q];
  }
  r.target_index = tbase + t;
  r.src_len = j.tl;
  r.rep_len = j.cl;
  r.pass = (uint8_t)pass;
  r.dead_flags = (uint8_t)(pass == 2 ? fp_flag_diff(tfp, cfp, t, slot) : 0);
  r.bytes_saved = (int16_t)((int)tbytes[t] - (int)cbytes[slot]);
  r.cycles_saved = (int16_t)((int)tcyc[t] - (int)ccyc[slot]);
  rules[o] = r;
}

// The empty candidate (the empty sequence: its outputs are its inputs, 0
// bytes, never pruned) against every target of a segment.  efp / emh are its
// fingerprint and Mid hashes (k_fingerprints of a length-0 sequence).  Go
// semantics: pass 1 when the exact QuickCheck (and MidCheck) hold, else pass
// 2 when dead flags are given and the masked ones hold; all hits (the CUDA
// tool): one masked pass.  Each such target gets an exhaustive job, planned
// here; slots are reserved once per wave.
__global__ __launch_bounds__(256) void k_empty_jobs(const uint2 *__restrict__ tfp, const uint2 *__restrict__ tmh,
                                                    const uint2 *__restrict__ efp, const uint2 *__restrict__ emh,
                                                    uint32_t nt, const uint32_t *__restrict__ targets, uint32_t tl,
                                                    uint32_t pm_masked, int hits_all, int dead_any, int use_mid,
                                                    int plan_mode, Job *jobs, uint32_t *njobs, uint32_t *chunks,
                                                    uint32_t *fail) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  uint32_t pass = 0;
  if (t < nt) {
    uint32_t acc_lo = 0, acc_hi = 0;
#pragma unroll
    for (int v = 0; v < 8; v++) {
      const uint2 a = tfp[(uint64_t)t * kFpVecs + v], b = efp[v];
      acc_lo |= a.x ^ b.x;
      acc_hi |= a.y ^ b.y;
    }
    const uint32_t mlo = 0x00FFFFFFu | (pm_masked << 24), flo = 0x00FFFFFFu | (kPMask << 24);
    const bool exact = ((acc_lo & flo) | acc_hi) == 0, masked = ((acc_lo & mlo) | acc_hi) == 0;
    const uint2 th = tmh[t], eh = emh[0];
    if (hits_all) pass = (masked && (!use_mid || th.y == eh.y)) ? 1u : 0u;
    else if (exact) pass = (!use_mid || th.x == eh.x) ? 1u : 0u;
    else if (dead_any && masked) pass = (!use_mid || th.y == eh.y) ? 2u : 0u;
  }
  const uint64_t m = __ballot(pass != 0);
  uint32_t base = 0;
  if (lane_id() == 0 && m) base = atomicAdd(njobs, (uint32_t)__popcll(m));
  base = uni(base);
  if (!pass) return;
  const uint32_t o = base + mbcnt(m);
  Job j{};
  for (uint32_t k = 0; k < tl; k++) j.tw[k] = targets[(uint64_t)t * tl + k];
  j.tl = (uint8_t)tl;
  j.cl = 0;
  j.pass = (uint8_t)pass;
  j.pmask = (uint8_t)((hits_all || pass == 2) ? pm_masked : kPMask);
  j.target = t;
  chunks[o] = plan_job(j, plan_mode);
  fail[o] = 0;
  jobs[o] = j;
}

// Rules of the empty-candidate jobs that passed; traw (length-3 ranges)
// maps a target's segment position to its raw triple offset.
__global__ __launch_bounds__(256) void k_empty_emit(const Job *jobs, uint32_t nj, const uint32_t *fail,
                                                    const uint2 *tfp, const uint2 *efp, const uint8_t *tbytes,
                                                    const uint8_t *tcyc, int hits_all, uint32_t dead, uint64_t tbase,
                                                    const uint32_t *traw, sx_rule *rules, uint32_t *nrules) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nj || fail[i]) return;
  const Job &j = jobs[i];
  uint32_t fd = 0;
  if (!hits_all && j.pass == 2) {
    for (int v = 0; v < 8; v++) fd |= ((tfp[(uint64_t)j.target * kFpVecs + v].x ^ efp[v].x) >> 24) & kPMask;
    if (fd == 0) return;  // FlagDiff == 0 is skipped (worker.go:211-213)
  }
  sx_rule r;
  memset(&r, 0, sizeof(r));
  for (int q = 0; q < 4; q++) r.src[q] = j.tw[q];
  r.target_index = tbase + (traw ? traw[j.target] : j.target);
  r.src_len = j.tl;
  r.rep_len = 0;
  r.pass = j.pass;
  r.dead_flags = (uint8_t)(hits_all ? dead : fd);
  r.bytes_saved = (int16_t)tbytes[j.target];
  r.cycles_saved = (int16_t)tcyc[j.target];
  rules[atomicAdd(nrules, 1u)] = r;
}

// ---- Candidates of length 3: a hash join over all 2,863^3 of them.
//
// A (target, pass) entry that no candidate of length 1 or 2 resolves goes on
// to the length-3 candidates (worker.go:122-123 loops candidate lengths in
// turn).  There are 23.5e9 of them, too many to list per target, so one sweep
// enumerates them on the fly for all open targets together: a wave holds 8
// prefixes (i0, i1) x the 8 QuickCheck vectors (lane = prefix * 8 + vector),
// runs each prefix once, then walks the 2,863 last instructions i2 in Go
// order with a wave-uniform opcode.  Each group of 8 lanes XORs its vectors'
// output hashes into the candidate's fingerprint hash, which is looked up in
// an LDS Bloom filter of the open targets' hashes and, when it passes, in
// their table in global memory.  Hits (the candidate, the table slot) are
// verified exactly against every target of the slot, filtered by bytes and
// the pruner and swept exhaustively; the first hit of an entry is the passing
// candidate with the smallest Go index (i0 * 2863 + i1) * 2863 + i2.
constexpr uint32_t kL3BloomLog = 18;                          // 2^18 bits = 32 KiB of LDS
constexpr uint32_t kL3BloomWords = (1u << kL3BloomLog) / 32;
constexpr uint32_t kL3MaxTargets = 8192;                      // open targets per sweep
constexpr uint32_t kL3Queue = 64;                             // LDS hit queue per wave
constexpr uint64_t kL3None = ~0ull;
constexpr uint64_t kL3Prefixes = (uint64_t)kNumInsts * kNumInsts;

// One vector's share of the fingerprint hash (P under the compare mask).
__device__ __forceinline__ uint32_t l3_mix(uint32_t lo, uint32_t hi, uint32_t v) {
  uint32_t x = (lo * 0x9E3779B1u) ^ hi ^ (v * 0x7FEB352Du);
  x ^= x >> 15;
  x *= 0x846CA68Bu;
  x ^= x >> 16;
  return x;
}
__device__ __forceinline__ uint32_t l3_bit1(uint32_t hsh) { return hsh & ((1u << kL3BloomLog) - 1u); }
__device__ __forceinline__ uint32_t l3_bit2(uint32_t hsh) { return (hsh * 0x5BD1E995u) >> (32 - kL3BloomLog); }

// Open targets: entry 2t (exact pass) without a rule, or with dead flags
// entry 2t + 1 without one, and more bytes than the shortest length-3
// candidate (3).  list <- t, hk <- its fingerprint hash (never 0);
// cnt[0] = how many, cnt[1] = their largest byte size.
__global__ __launch_bounds__(256) void k_l3_open(const uint32_t *rule_flag, uint32_t nt, int dead_any,
                                                 const uint8_t *tbytes, const uint2 *tfp, uint32_t mlo,
                                                 uint32_t *list, uint32_t *hk, uint32_t *cnt) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  bool open = false;
  if (t < nt) open = tbytes[t] > 3 && (!rule_flag[2 * t] || (dead_any && !rule_flag[2 * t + 1]));
  const uint64_t m = __ballot(open);
  if (m == 0) return;
  uint32_t base = 0;
  if (lane_id() == 0) base = atomicAdd(&cnt[0], (uint32_t)__popcll(m));
  base = __shfl(base, 0);
  if (!open) return;
  uint32_t hsh = 0;
  for (uint32_t v = 0; v < kFpVecs; v++) {
    const uint2 f = tfp[(uint64_t)t * kFpVecs + v];
    hsh ^= l3_mix(f.x & mlo, f.y, v);
  }
  const uint32_t o = base + mbcnt(m);
  list[o] = t;
  hk[o] = hsh | 1u;
  atomicMax(&cnt[1], (uint32_t)tbytes[t]);
}

// The table of n open targets' hashes: open addressing over mask + 1 >= 4n
// slots (keys: 0 = free), each slot heading a list of the targets with that
// hash (head -1 = none, next); and the Bloom filter's two bits per hash.
__global__ __launch_bounds__(256) void k_l3_insert(const uint32_t *hk, uint32_t n, uint32_t mask, uint32_t *keys,
                                                   int32_t *head, int32_t *next, uint32_t *bloom) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint32_t hsh = hk[i];
  uint32_t sl = hsh & mask;
  for (;;) {  // more than n slots are free, so a free or matching slot is reached
    const uint32_t old = atomicCAS(&keys[sl], 0u, hsh);
    if (old == 0u || old == hsh) break;
    sl = (sl + 1) & mask;
  }
  next[i] = atomicExch(&head[sl], (int32_t)i);
  atomicOr(&bloom[l3_bit1(hsh) >> 5], 1u << (l3_bit1(hsh) & 31u));
  atomicOr(&bloom[l3_bit2(hsh) >> 5], 1u << (l3_bit2(hsh) & 31u));
}

// The sweep over prefixes [p0, p0 + np): raw hits (prefix, i2 | slot << 12).
__global__ __launch_bounds__(256) void k_l3_join(uint64_t p0, uint64_t np, const uint32_t *__restrict__ keys,
                                                 uint32_t mask, const uint32_t *__restrict__ bloom, int max_tbytes,
                                                 int model, uint32_t mlo, uint2 *__restrict__ raw, uint32_t cap,
                                                 uint32_t *__restrict__ count) {
  __shared__ uint32_t s_bloom[kL3BloomWords];
  __shared__ uint2 queue[4][kL3Queue];
  for (uint32_t i = threadIdx.x; i < kL3BloomWords; i += 256) s_bloom[i] = bloom[i];
  __syncthreads();
  const uint32_t lane = lane_id(), wv = uni(threadIdx.x >> 6);
  const uint32_t g = lane >> 3, v = lane & 7;
  const uint64_t p = p0 + ((uint64_t)blockIdx.x * 4 + wv) * 8 + g;
  bool active = p < p0 + np;
  const uint32_t pi = active ? (uint32_t)p : 0u;
  uint32_t w[2] = {go_inst(pi / kNumInsts), go_inst(pi % kNumInsts)};
  const int pbytes = byte_size(w[0] & 0xFFFFu, model) + byte_size(w[1] & 0xFFFFu, model);
  // a prunable prefix makes every extension prunable (the redundancy inside
  // it stays in the candidate)
  if (should_prune(w, 2, c_reads, c_writes)) active = false;
  if (pbytes + 1 >= max_tbytes) active = false;
  State s;
  s.a = c_vec[v][0]; s.x = c_vec[v][1]; s.y = c_vec[v][2]; s.p = c_vec[v][3];
  s.s = c_vec[v][4]; s.m = c_vec[v][5]; s.s0 = c_vec[v][6]; s.s1 = c_vec[v][7];
  step(s, w[0] & 0xFFFFu, w[0] >> 16);
  step(s, w[1] & 0xFFFFu, w[1] >> 16);
  uint32_t qn = 0;
  auto flush = [&]() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(count, qn);
    base = uni(base);
    for (uint32_t i = lane; i < qn; i += kWave)
      if (base + i < cap) raw[base + i] = queue[wv][i];
    __builtin_amdgcn_wave_barrier();
    qn = 0;
  };
  if (__ballot(active) != 0) {
    for (uint32_t i2 = 0; i2 < (uint32_t)kNumInsts; i2++) {
      const uint32_t wi = go_inst(i2), op = wi & 0xFFFFu;
      // Go order: the one-byte instructions first, then the two-byte ones
      if (i2 == (uint32_t)kNumNonImm && __ballot(active && pbytes + 2 < max_tbytes) == 0) break;
      const bool ok_b = active && pbytes + byte_size(op, model) < max_tbytes;
      State s2 = s;
      step(s2, op, wi >> 16);
      uint32_t hsh = l3_mix(pack_lo(s2) & mlo, pack_hi(s2), v);
      hsh ^= __shfl_xor(hsh, 1);
      hsh ^= __shfl_xor(hsh, 2);
      hsh ^= __shfl_xor(hsh, 4);
      hsh |= 1u;
      bool cand = false;
      if (ok_b && v == 0) {
        const uint32_t b1 = l3_bit1(hsh), b2 = l3_bit2(hsh);
        cand = ((s_bloom[b1 >> 5] >> (b1 & 31u)) & (s_bloom[b2 >> 5] >> (b2 & 31u)) & 1u) != 0;
      }
      if (__ballot(cand) == 0) continue;
      uint32_t slot = 0;
      bool hit = false;
      if (cand) {
        uint32_t sl = hsh & mask;
        for (;;) {  // the table has free slots, so the probe ends
          const uint32_t k = keys[sl];
          if (k == hsh) {
            hit = true;
            slot = sl;
            break;
          }
          if (k == 0u) break;
          sl = (sl + 1) & mask;
        }
      }
      const uint64_t hm = __ballot(hit);
      if (hm == 0) continue;
      if (hit) queue[wv][qn + (uint32_t)__popcll(hm & ((1ull << lane) - 1))] = make_uint2(pi, i2 | (slot << 12));
      qn += (uint32_t)__popcll(hm);
      if (qn > kL3Queue - 8) flush();
    }
  }
  if (qn) flush();
}

// Outputs of a length-3 candidate on the 8 QuickCheck vectors.
__device__ __forceinline__ void l3_outputs(const uint32_t w[3], uint32_t lo[8], uint32_t hi[8]) {
  for (uint32_t v = 0; v < kFpVecs; v++) {
    State s;
    s.a = c_vec[v][0]; s.x = c_vec[v][1]; s.y = c_vec[v][2]; s.p = c_vec[v][3];
    s.s = c_vec[v][4]; s.m = c_vec[v][5]; s.s0 = c_vec[v][6]; s.s1 = c_vec[v][7];
    for (int k = 0; k < 3; k++) step(s, w[k] & 0xFFFFu, w[k] >> 16);
    lo[v] = pack_lo(s);
    hi[v] = pack_hi(s);
  }
}
__device__ __forceinline__ void l3_words(uint64_t c3, uint32_t w[3]) {
  const uint32_t p = (uint32_t)(c3 / kNumInsts), i2 = (uint32_t)(c3 % kNumInsts);
  w[0] = go_inst(p / kNumInsts);
  w[1] = go_inst(p % kNumInsts);
  w[2] = go_inst(i2);
}

// Raw hits -> exhaustive jobs: every target of the hit's slot (list[] maps
// the sweep's entries to segment positions) whose QuickCheck the candidate
// passes exactly (pass 1, entry open) or only under the mask (pass 2, dead
// flags, entry open), with fewer bytes, the candidate not pruned.
__global__ __launch_bounds__(256) void k_l3_expand(const uint2 *raw, uint32_t nraw, const int32_t *head,
                                                   const int32_t *next, const uint32_t *list,
                                                   const uint32_t *targets, uint32_t tl, const uint2 *tfp,
                                                   const uint8_t *tbytes, const uint32_t *rule_flag, int dead_any,
                                                   uint32_t pm_masked, int model, Job *jobs, uint64_t *c3out,
                                                   uint32_t cap, uint32_t *njobs) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nraw) return;
  const uint2 e = raw[i];
  const uint64_t c3 = (uint64_t)e.x * kNumInsts + (e.y & 0xFFFu);
  uint32_t w[3];
  l3_words(c3, w);
  if (should_prune(w, 3, c_reads, c_writes)) return;
  int cb = 0;
  for (int k = 0; k < 3; k++) cb += byte_size(w[k] & 0xFFFFu, model);
  uint32_t lo[8], hi[8];
  l3_outputs(w, lo, hi);
  const uint32_t mlo = 0x00FFFFFFu | (pm_masked << 24), flo = 0x00FFFFFFu | (kPMask << 24);
  for (int32_t q = head[e.y >> 12]; q >= 0; q = next[q]) {
    const uint32_t t = list[q];
    if (cb >= (int)tbytes[t]) continue;
    uint32_t dl = 0, dh = 0;
    for (uint32_t v = 0; v < kFpVecs; v++) {
      const uint2 f = tfp[(uint64_t)t * kFpVecs + v];
      dl |= f.x ^ lo[v];
      dh |= f.y ^ hi[v];
    }
    const bool okf = ((dl & flo) | dh) == 0, okm = ((dl & mlo) | dh) == 0;
    uint32_t pass = 0;
    if (okf && !rule_flag[2 * t]) pass = 1;
    else if (!okf && okm && dead_any && !rule_flag[2 * t + 1]) pass = 2;
    if (!pass) continue;
    const uint32_t o = atomicAdd(njobs, 1u);
    if (o >= cap) continue;
    Job j;
    for (int k = 0; k < kJobWords; k++) {
      j.tw[k] = k < (int)tl ? targets[(uint64_t)t * tl + k] : 0u;
      j.cw[k] = k < 3 ? w[k] : 0u;
    }
    j.tl = (uint8_t)tl;
    j.cl = 3;
    j.pass = (uint8_t)pass;
    j.pmask = (uint8_t)(pass == 2 ? pm_masked : kPMask);
    j.target = t;
    j.slot = 0;
    jobs[o] = j;
    c3out[o] = c3;
  }
}

// FlagDiff (verifier.go:591-604) of a target and a length-3 candidate from
// the target's fingerprint; 0 when a non-P byte differs.
__device__ __forceinline__ uint32_t l3_flag_diff(const uint2 *tfp, uint32_t t, const uint32_t w[3]) {
  uint32_t lo[8], hi[8];
  l3_outputs(w, lo, hi);
  uint32_t fd = 0;
  for (uint32_t v = 0; v < kFpVecs; v++) {
    const uint2 f = tfp[(uint64_t)t * kFpVecs + v];
    if (((f.x ^ lo[v]) & 0x00FFFFFFu) || (f.y ^ hi[v])) return 0;
    fd |= ((f.x ^ lo[v]) >> 24) & kPMask;
  }
  return fd;
}

// The passing jobs' smallest Go index per (target, pass); a flag-relaxed
// rule needs FlagDiff != 0 (worker.go:211-213).
__global__ __launch_bounds__(256) void k_l3_best(const Job *jobs, uint32_t nj, const uint32_t *fail,
                                                 const uint64_t *c3, const uint2 *tfp, unsigned long long *best) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nj || fail[i]) return;
  const Job &j = jobs[i];
  if (j.pass == 2) {
    uint32_t w[3] = {j.cw[0], j.cw[1], j.cw[2]};
    if (l3_flag_diff(tfp, j.target, w) == 0) return;
  }
  atomicMin(&best[(uint64_t)j.target * 2 + (j.pass - 1)], (unsigned long long)c3[i]);
}

// Rules of the entries the length-3 candidates resolved, in their slots.
__global__ __launch_bounds__(256) void k_l3_emit(const unsigned long long *best, uint32_t nt, const uint32_t *targets,
                                                 uint32_t tl, const uint2 *tfp, const uint8_t *tbytes,
                                                 const uint8_t *tcyc, int model, uint64_t tbase, sx_rule *rules,
                                                 uint32_t *rule_flag) {
  const uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= 2 * nt || best[k] == kL3None) return;
  const uint32_t t = k >> 1, pass = (k & 1) + 1;
  uint32_t w[3];
  l3_words(best[k], w);
  sx_rule r;
  memset(&r, 0, sizeof(r));
  int cb = 0, cc = 0;
  for (int q = 0; q < 4; q++) {
    r.src[q] = q < (int)tl ? targets[(uint64_t)t * tl + q] : 0u;
    r.rep[q] = q < 3 ? w[q] : 0u;
  }
  for (int q = 0; q < 3; q++) {
    cb += byte_size(w[q] & 0xFFFFu, model);
    cc += cycles(w[q] & 0xFFFFu);
  }
  r.target_index = tbase + t;
  r.src_len = (uint8_t)tl;
  r.rep_len = 3;
  r.pass = (uint8_t)pass;
  r.dead_flags = (uint8_t)(pass == 2 ? l3_flag_diff(tfp, t, w) : 0u);
  r.bytes_saved = (int16_t)((int)tbytes[t] - cb);
  r.cycles_saved = (int16_t)((int)tcyc[t] - cc);
  rules[k] = r;
  rule_flag[k] = 1;
}

// Length-3 ranges: a rule's target_index (base + position in the segment)
// becomes base + the target's offset in the raw triple index range.
__global__ void k_fix_raw_index(sx_rule *rules, uint32_t n, const uint32_t *raw_off, uint64_t base) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) rules[i].target_index = base + raw_off[rules[i].target_index - base];
}

// Ordered compaction of the first-hit rule slots (flag[k] != 0 where slot k
// holds a rule), tile by tile as in the scan: sums = the exclusive scan of
// the tiles' flag counts (k_tile_sums with flags, then k_scan_sums).  The
// tile's filled slots are listed in LDS in order, then the block copies the
// listed records with consecutive threads (one record per thread and pass),
// so the copies of a tile are all in flight at once.
__global__ __launch_bounds__(256) void k_compact_rules(const sx_rule *slots, const uint32_t *flag, uint64_t n,
                                                       const uint64_t *sums, sx_rule *out) {
  __shared__ uint64_t s_w[4];
  __shared__ uint16_t s_idx[kScanTile];
  const uint32_t t0 = threadIdx.x * kScanItems;
  const uint64_t tile = (uint64_t)blockIdx.x * kScanTile;
  uint32_t v[kScanItems];
  scan_load(flag, n, tile + t0, v);
  uint32_t t = 0;
#pragma unroll
  for (int i = 0; i < (int)kScanItems; i++) t += v[i] != 0u;
  uint64_t tot = 0;
  uint32_t o = (uint32_t)block_scan256(t, s_w, &tot);
#pragma unroll
  for (int i = 0; i < (int)kScanItems; i++)
    if (v[i]) s_idx[o++] = (uint16_t)(t0 + i);
  __syncthreads();
  const uint64_t ob = sums[blockIdx.x];
  for (uint32_t j = threadIdx.x; j < (uint32_t)tot; j += 256) out[ob + j] = slots[tile + s_idx[j]];
}

// Rules of length-2 targets against single-instruction candidates packed
// for the trip to the host: 3 words instead of the 12 of an sx_rule.  An
// instruction word (opcode below 64, immediate below 256 at bit 16) takes 14
// bits as opcode | immediate << 6: word 0 = src[0] | src[1] << 14 | pass <<
// 28, word 1 = rep[0] | dead flags << 14, word 2 = the target index (below
// 2^32).  The lengths are 2 and 1 and the savings follow from the words, so
// the host fills them in.  The copy over the host link is the one part of an
// enumeration that its own compute cannot hide.
constexpr uint32_t kPackedWords = 3;
static_assert(sizeof(sx_rule) == 48, "sx_rule is 12 words");
__host__ __device__ inline uint32_t inst14(uint32_t w) { return (w & 63u) | ((w >> 16) & 0xFFu) << 6; }
__host__ __device__ inline uint32_t inst_of14(uint32_t v) { return (v & 63u) | ((v >> 6) & 0xFFu) << 16; }
__global__ __launch_bounds__(256) void k_pack_rules(const sx_rule *r, uint32_t n, uint32_t *w) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const sx_rule x = r[i];
  w[3 * i] = inst14(x.src[0]) | inst14(x.src[1]) << 14 | (uint32_t)x.pass << 28;
  w[3 * i + 1] = inst14(x.rep[0]) | (uint32_t)x.dead_flags << 14;
  w[3 * i + 2] = (uint32_t)x.target_index;
}

// Bulk QuickCheck / MidCheck / FlagDiff for sx_check_pairs: one wave per pair.
__global__ __launch_bounds__(256) void k_pair_vectors(const uint32_t *tw, const uint8_t *tl,
                                                      const uint32_t *cw, const uint8_t *cl,
                                                      uint64_t n, uint32_t pm, uint8_t *qc,
                                                      uint8_t *mid, uint8_t *fdiff) {
  uint64_t p = (uint64_t)blockIdx.x * 4 + uni(threadIdx.x >> 6);  // wave-uniform
  if (p >= n) return;
  uint32_t lane = lane_id();
  uint32_t t[4], c[4];
  uint32_t tn = uni(tl[p]), cn = uni(cl[p]);
  for (int k = 0; k < 4; k++) {
    t[k] = uni(tw[p * 4 + k]);
    c[k] = uni(cw[p * 4 + k]);
  }
  bool eq = true, nonp = true;
  uint32_t fd = 0;
  if (lane < 32) {
    State a = load_vec(lane), b = a;
    run_seq_uniform(a, t, (int)tn);
    run_seq_uniform(b, c, (int)cn);
    nonp = ((a.a ^ b.a) | (a.x ^ b.x) | (a.y ^ b.y) | (a.s ^ b.s) | (a.m ^ b.m) | (a.s0 ^ b.s0) |
            (a.s1 ^ b.s1)) == 0;
    eq = nonp && ((a.p ^ b.p) & pm) == 0;
    fd = (a.p ^ b.p) & kPMask;
  }
  uint64_t bad_all = __ballot(!eq);
  uint64_t bad_qc = bad_all & 0xFFull;
  uint64_t nonp_bad = __ballot(!nonp) & 0xFFull;
  // FlagDiff over the 8 QuickCheck vectors (verifier.go:591-604)
  uint32_t f = lane < 8 ? fd : 0;
  for (int off = 32; off >= 1; off >>= 1) f |= __shfl_xor(f, off);
  if (lane == 0) {
    if (qc) qc[p] = bad_qc == 0;
    if (mid) mid[p] = bad_all == 0;
    if (fdiff) fdiff[p] = nonp_bad ? 0 : (uint8_t)f;
  }
}

// Exhaustive jobs of sx_check_pairs' pairs built on the device (4 words per
// sequence; words beyond a sequence's length are zeroed).
__global__ void k_pair_jobs(const uint32_t *tw, const uint8_t *tl, const uint32_t *cw, const uint8_t *cl, uint32_t n,
                            uint32_t pm, Job *jobs) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Job j{};
  const uint32_t a = tl[i], b = cl[i];
  for (uint32_t k = 0; k < 4; k++) {
    j.tw[k] = k < a ? tw[(uint64_t)i * 4 + k] : 0u;
    j.cw[k] = k < b ? cw[(uint64_t)i * 4 + k] : 0u;
  }
  j.tl = (uint8_t)a;
  j.cl = (uint8_t)b;
  j.pass = 1;
  j.pmask = (uint8_t)pm;
  j.target = i;
  jobs[i] = j;
}

// sx_check_pairs' per-pair result bits from the stage outputs (qc, mid: 1 =
// holds; fail: 0 = the exhaustive check passed); null = stage not run.
__global__ void k_pair_bits(const uint8_t *qc, const uint8_t *mid, const uint32_t *fail, uint32_t n,
                            uint32_t stage_mask, uint8_t *out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t v = 0;
  if ((stage_mask & SX_STAGE_QC) && qc && qc[i]) v |= SX_STAGE_QC;
  if ((stage_mask & SX_STAGE_MID) && mid && mid[i]) v |= SX_STAGE_MID;
  if ((stage_mask & SX_STAGE_EXH) && fail && !fail[i]) v |= SX_STAGE_EXH;
  out[i] = (uint8_t)v;
}

__global__ void k_exec_states(uint8_t *states, const uint32_t *words, const uint8_t *len, uint64_t n,
                              uint32_t *bad_op) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint8_t *b = states + i * 8;
  State s;
  s.a = b[0]; s.x = b[1]; s.y = b[2]; s.p = b[3]; s.s = b[4]; s.m = b[5]; s.s0 = b[6]; s.s1 = b[7];
  for (int k = 0; k < len[i]; k++) {
    uint32_t w = words[i * 4 + k];
    if (!step(s, w & 0xFFFFu, w >> 16)) atomicOr(bad_op, 1u);
  }
  b[0] = s.a; b[1] = s.x; b[2] = s.y; b[3] = s.p; b[4] = s.s; b[5] = s.m; b[6] = s.s0; b[7] = s.s1;
}

__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_fill_u8(uint8_t *p, uint64_t n, uint8_t v) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Candidate ranks: slot order is Go order, so the rank is the slot.
__global__ void k_iota(uint32_t *p, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (uint32_t)i;
}

// The 2,863 instruction words in Go order, one word each.
__global__ void k_inst_words(uint32_t *w) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < (uint32_t)kNumInsts) w[i] = go_inst(i);
}

__global__ void k_cand_words(uint32_t *w, uint64_t n) {
  // 2 words per slot: single instructions then Go-ordered pairs
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (i < (uint64_t)kNumInsts) {
    w[2 * i] = go_inst((uint32_t)i);
    w[2 * i + 1] = 0;
  } else {
    uint64_t s = i - kNumInsts;
    w[2 * i] = go_inst((uint32_t)(s / kNumInsts));
    w[2 * i + 1] = go_inst((uint32_t)(s % kNumInsts));
  }
}

// -------------------------------------------------------------- host side --

namespace {

struct DevBuf {
  void *p = nullptr;
  size_t cap = 0;
};

}  // namespace

struct sx_handle {
  int device = 0;
  int num_cu = 256;
  bool poisoned = false;  // a device stage timed out: the stream is unusable
  int sort_bits = 10;     // key bits per pass of the group sort (SX_SORT_BITS: 8 or 10)
  int bulk_after = -2;    // first-hit rounds before the one-batch tail (SX_BULK_AFTER); -1 = by survivor count, -2 = 0 or 1 by survivor count
  int exh_k = 2;          // states (pairs when packed) per lane and opcode dispatch in the sweep (SX_EXH_K: 1 or 2)
  int exh_pk = 1;         // packed two-states-per-register interpreter in the sweep (SX_EXH_PK=0: one state)
  int qcm_join = 1;       // QuickCheck and MidCheck as one join when MidCheck is on (SX_QCM_JOIN=0: separate)
  int fp_split = 1;       // the join's targets: QuickCheck outputs first, Mid hashes of class members (SX_FP_SPLIT)
  int fetch_threads = 16;  // host threads of sx_fetch_rules (SX_FETCH_THREADS, 1..64)
  // sx_fetch_rules' helper threads, started on its first call and kept (a
  // thread costs tens of microseconds to start, 15 of them more than the
  // unpacking): each call posts its work and runs it on the caller's thread
  // too, then waits until every helper has left it
  struct FetchPool {
    std::vector<std::thread> th;
    std::mutex mu;
    std::condition_variable cv, cv_done;
    uint64_t gen = 0;
    const std::function<void()> *job = nullptr;
    size_t busy = 0;
    bool stop = false;
  } fetch_pool;
  // the candidate side (words, fingerprints, hash tables, classes) is kept
  // between enumerations while the candidate length, byte model and compare
  // mask stay the same
  struct {
    bool valid = false, fused_built = false, fused_ok = false;
    int ml = 0, bytes_model = -1;
    uint32_t lo_mask = 0;
  } cand;
  uint32_t mid_cap_hint = 0;  // MidCheck survivor buffer size that sufficed before
  int exh_group = 4;      // small plans swept one job per lane, this many pairs per lane and dispatch
                          // (SX_EXH_GROUP: 1, 2, 4 or 8; 0: every job state-parallel)
  // the plan table in B_PLANTAB holds the plans of (sweep mode, relaxed mask)
  // = (ptab_mode, ptab_pm); -1: none yet
  int ptab_mode = -1;
  uint32_t ptab_pm = 0;
  int seg_log = 23;       // targets per segment = 2^seg_log (SX_SEG_LOG, 16..24)
  uint32_t surv_cap_hint = 0;  // QuickCheck survivor buffer size that sufficed before
  // kernel-only timing of the QuickCheck and exhaustive-sweep launches
  hipEvent_t kev[2] = {nullptr, nullptr};
  double k_qc_ms = 0, k_exh_ms = 0, k_grp_ms = 0;
  uint64_t k_qc_launches = 0, k_exh_launches = 0, k_grp_launches = 0;
  // event pairs around QuickCheck and sweep launches, read at the next host
  // synchronisation instead of synchronising per launch
  struct Timed {
    hipEvent_t a, b;
    int kind;  // 0 QuickCheck, 1 exhaustive sweep
  };
  std::vector<Timed> timed;
  size_t timed_used = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::vector<sx_rule> rules;  // all-hits mode: sorted on the host
  std::vector<sx_rule> empty_rules;  // the empty-candidate stream (include_empty), sorted on the host
  // first-hit mode: the rules stay on the device, in order, until fetched,
  // rec_words 32-bit words per rule: 12 (an sx_rule) or 3 (the packed form of
  // k_pack_rules, which sx_fetch_rules expands)
  uint32_t *d_rules = nullptr;
  uint64_t d_rules_capw = 0, d_rules_n = 0;
  uint32_t rec_words = 12;
  bool rules_on_device = false;
  // ... and a pinned host copy of them, filled segment by segment on
  // copy_stream while the later segments compute
  hipStream_t copy_stream = nullptr;
  hipStream_t copy_stream2 = nullptr;  // the odd pieces' (SX_COPY_STREAMS=2, the default)
  int copy_streams = 2;
  hipEvent_t seg_ev = nullptr;
  // the host copy arrives in pieces of kCopyPiece rules, each with an event
  // on copy_stream, so that sx_fetch_rules unpacks the first pieces while the
  // later ones are still on the link; pieces[i] = (first rule, count)
  std::vector<hipEvent_t> piece_ev;
  std::vector<std::pair<uint64_t, uint64_t>> pieces;
  uint32_t *h_rules = nullptr;
  uint64_t h_rules_capw = 0;
  int fetch_model = 0;  // the byte model of the enumeration whose rules are held
  // pinned staging block for the counters the host reads back to size the
  // next stage (a copy to pageable memory waits by itself, so two such reads
  // cost two round trips)
  uint8_t *pin = nullptr;
  // QuickCheck launch configurations chosen by the autotuner, per (target_len, max_cand_len)
  struct QcChoice {
    int tl, ml, big;  // big: segments of >= 65,536 targets
    uint32_t tile, blocks_per_cu;  // blocks_per_cu 0 = grid form; tile 0 too = the hash join
  };
  // the hash join's candidate side for the current enumeration (sorted filter
  // hashes, their slots, the bucket directory on the top qj_dk bits)
  const uint32_t *qj_hash = nullptr, *qj_slot = nullptr, *qj_dir = nullptr;
  uint32_t qj_dk = 1;
  std::vector<QcChoice> qc_tuned;
  // cached pruned length-2 target list (Go order)
  uint32_t *d_l2 = nullptr;
  uint64_t n_l2 = 0;
  // buffers reused across calls
  DevBuf bufs[96];
  // device address of the no-alloc scratch array on this handle's device
  // (null when the lookup failed; scratch_rc says why)
  uint8_t *scratch = nullptr;
  hipError_t scratch_rc = hipSuccess;
};

namespace {

enum BufId {
  B_TARGETS, B_TFP, B_TV0, B_TBYTES, B_TCYC, B_CWORDS, B_CFP, B_CBYTES, B_CCYC, B_CPRUNE,
  B_RANK, B_SURV, B_MID, B_COUNTERS, B_CUR, B_REJ, B_DONE, B_JOBS, B_JENT, B_CHUNKS, B_OFFS, B_SUMS,
  B_FAIL, B_RULES, B_SCRATCH, B_GEN_COUNTS, B_GEN_OFFS, B_BADOP, B_P_TW, B_P_TL, B_P_CW, B_P_CL,
  B_P_OUT1, B_P_OUT2, B_P_OUT3, B_STATES, B_TOTAL, B_CURSOR, B_HEADS, B_RULEFLAG, B_RULES2,
  B_RULETOT, B_TRAW, B_CHUNKMAP, B_TMH, B_CMH, B_PLANTAB, B_EFP, B_ERULES, B_L3_LIST, B_L3_HK, B_L3_CNT,
  B_L3_KEYS, B_L3_HEAD, B_L3_NEXT, B_L3_BLOOM, B_L3_RAW, B_L3_C3, B_L3_BEST, B_GKEY, B_GIDX, B_GKEY_S, B_GIDX_S,
  B_GSORT, B_QJ_KEY, B_QJ_VAL, B_QJ_HASH, B_QJ_SLOT, B_QJ_DIR, B_CK, B_CKS, B_CV1,
  B_CV1S, B_JK, B_JV, B_JKS, B_JVS, B_JDIR, B_CHEAD, B_CHOFF, B_CLSKEY, B_CLSREP, B_CLSCNT, B_CLSN, B_QCSUM, B_TLIST, B_EVSUM, B_CLSDIR, B_TCLS, B_CCLS, B_CPSIG, B_MIDCNT, B_COUNT
};
static_assert(B_COUNT <= 96, "sx_handle::bufs holds every buffer id");

int set_err(sx_handle *h, int code, const char *fmt, ...) {
  if (h) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    h->err = buf;
  }
  return code;
}

#define HIPCHK(h, call)                                                                        \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      return set_err((h), SX_EHIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),      \
                     __FILE__, __LINE__);                                                      \
  } while (0)

// Stream synchronisation with a deadline: a stage that does not finish within
// SX_WATCHDOG_S seconds (default 120) fails with SX_EHIP instead of blocking
// the caller forever.  The handle must not be used after such an error.
// Kernel timing: record the start of a timed launch (returns the pool slot)
// and its end; the durations are added up by harvest_timed once the stream
// has drained.
int timed_begin(sx_handle *h, int kind, size_t *slot) {
  if (h->timed_used == h->timed.size()) {
    sx_handle::Timed t{nullptr, nullptr, 0};
    HIPCHK(h, hipEventCreate(&t.a));
    if (hipEventCreate(&t.b) != hipSuccess) {
      (void)hipEventDestroy(t.a);
      return set_err(h, SX_EHIP, "hipEventCreate failed");
    }
    h->timed.push_back(t);
  }
  *slot = h->timed_used++;
  h->timed[*slot].kind = kind;
  HIPCHK(h, hipEventRecord(h->timed[*slot].a, h->stream));
  return SX_OK;
}
int timed_end(sx_handle *h, size_t slot) {
  HIPCHK(h, hipEventRecord(h->timed[slot].b, h->stream));
  return SX_OK;
}
void harvest_timed(sx_handle *h) {
  for (size_t i = 0; i < h->timed_used; i++) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, h->timed[i].a, h->timed[i].b) != hipSuccess) {
      (void)hipGetLastError();
      continue;
    }
    if (h->timed[i].kind == 0) {
      h->k_qc_ms += ms;
      h->k_qc_launches++;
    } else if (h->timed[i].kind == 2) {
      h->k_grp_ms += ms;
      h->k_grp_launches++;
    } else {
      h->k_exh_ms += ms;
      h->k_exh_launches++;
    }
  }
  h->timed_used = 0;
}

int sync_wd(sx_handle *h, const char *stage) {
  static double limit = [] {
    const char *e = getenv("SX_WATCHDOG_S");
    return e ? atof(e) : 120.0;
  }();
  auto t0 = std::chrono::steady_clock::now();
  for (;;) {
    hipError_t e = hipStreamQuery(h->stream);
    if (e == hipSuccess) {
      harvest_timed(h);
      return SX_OK;
    }
    if (e != hipErrorNotReady)
      return set_err(h, SX_EHIP, "%s: %s", stage, hipGetErrorString(e));
    double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (dt > limit) {
      h->poisoned = true;
      return set_err(h, SX_EHIP, "%s: device work did not finish within %.0f s", stage, limit);
    }
    // Most waits here are a few kernels (tens of microseconds to a few
    // milliseconds: the length-2 sweep is one 2 ms launch, a 0xFF join 7 ms):
    // spin for the first 20 ms (a timed sleep oversleeps by the timer slack,
    // 50-100 us, on every wait), then poll at 50 us and after 0.1 s at 0.5 ms.
    if (dt < 0.02) std::this_thread::yield();
    else std::this_thread::sleep_for(std::chrono::microseconds(dt < 0.1 ? 50 : 500));
  }
}

// A segment's packed rules whose copy to the host is yet to be queued: the
// words at offset used of the device and host lists, n rules from rule
// first on, rw words each.
struct PendingRules {
  uint64_t used, n, first, rw;
};
// Queue the copy of p's rules in pieces of kCopyPiece rules, each with an
// event that sx_fetch_rules polls; the copy stream already waits for the
// packing.
int queue_pieces(sx_handle *h, PendingRules &p) {
  for (uint64_t r0 = 0; r0 < p.n; r0 += kCopyPiece) {
    const uint64_t pn = std::min<uint64_t>(kCopyPiece, p.n - r0);
    const size_t pi = h->pieces.size();
    // alternate pieces over two copy streams (two copy engines)
    hipStream_t cs = (h->copy_streams == 2 && (pi & 1)) ? h->copy_stream2 : h->copy_stream;
    if (pi == h->piece_ev.size()) {
      hipEvent_t ev;
      HIPCHK(h, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
      h->piece_ev.push_back(ev);
    }
    HIPCHK(h, hipMemcpyAsync(h->h_rules + p.used + r0 * p.rw, h->d_rules + p.used + r0 * p.rw,
                             (size_t)pn * p.rw * 4, hipMemcpyDeviceToHost, cs));
    HIPCHK(h, hipEventRecord(h->piece_ev[pi], cs));
    h->pieces.emplace_back(p.first + r0, pn);
  }
  p.n = 0;
  return SX_OK;
}

// Device-to-host reads of small counters through the handle's pinned staging
// block: the copies are queued together and waited for once.
struct ReadBack {
  void *dst;
  const void *src;
  size_t bytes;
};
// pend: rule copies queued between the read-back's copies and the wait.
int read_back(sx_handle *h, std::initializer_list<ReadBack> rs, const char *stage, PendingRules *pend = nullptr) {
  size_t off = 0;
  for (const ReadBack &r : rs) {
    if (off + r.bytes > kPinBytes) return set_err(h, SX_EINVAL, "%s: read-back of %zu bytes", stage, off + r.bytes);
    HIPCHK(h, hipMemcpyAsync(h->pin + off, r.src, r.bytes, hipMemcpyDeviceToHost, h->stream));
    off += (r.bytes + 63) & ~size_t(63);
  }
  if (pend && pend->n) {
    int rcq = queue_pieces(h, *pend);
    if (rcq) return rcq;
  }
  int rc = sync_wd(h, stage);
  if (rc) return rc;
  off = 0;
  for (const ReadBack &r : rs) {
    memcpy(r.dst, h->pin + off, r.bytes);
    off += (r.bytes + 63) & ~size_t(63);
  }
  return SX_OK;
}

// SX_TRACE=1 prints stage progress to stderr.
void trace(sx_handle *h, const char *fmt, ...) {
  static int on = getenv("SX_TRACE") ? atoi(getenv("SX_TRACE")) : 0;
  if (!on) return;
  (void)h;
  va_list ap;
  va_start(ap, fmt);
  fprintf(stderr, "[sx] ");
  vfprintf(stderr, fmt, ap);
  fprintf(stderr, "\n");
  fflush(stderr);
  va_end(ap);
}

template <typename T>
int ensure(sx_handle *h, BufId id, size_t count, T **out) {
  DevBuf &b = h->bufs[id];
  size_t bytes = std::max<size_t>(count * sizeof(T), 256);
  if (b.cap < bytes) {
    if (b.p) {
      int rcw = sync_wd(h, "buffer resize");
      if (rcw) return rcw;
      HIPCHK(h, hipFree(b.p));
      b.p = nullptr;
      b.cap = 0;
    }
    size_t alloc = bytes + bytes / 8;
    if (hipMalloc(&b.p, alloc) != hipSuccess) {
      (void)hipGetLastError();
      return set_err(h, SX_ENOMEM, "hipMalloc(%zu) failed", alloc);
    }
    b.cap = alloc;
  }
  *out = (T *)b.p;
  return SX_OK;
}

#define SYNC_WD(stage)              \
  do {                              \
    int rcw_ = sync_wd(h, (stage)); \
    if (rcw_) return rcw_;          \
  } while (0)

#define ENSURE(id, T, n, ptr)                        \
  do {                                               \
    int rc_ = ensure<T>(h, (id), (size_t)(n), &(ptr)); \
    if (rc_) return rc_;                             \
  } while (0)

inline unsigned blocks_for(uint64_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

// The packed rules of k_pack_rules back into sx_rule records (model: the
// byte model of the enumeration): words 0-1 src[0..1], 4 rep[0], 8 the target
// index, 10 the lengths (2, 1), dead flags and pass, 11 the bytes and cycles
// saved, from per-opcode tables; the rest 0.  Streaming (non-temporal)
// 16-byte stores when out is 16-byte aligned: the records are written once
// and not read back here, and such stores skip reading the destination lines
// first.
void unpack_rules(sx_rule *out, const uint32_t *w, uint64_t n, int model) {
  static_assert(sizeof(sx_rule) == 48 && offsetof(sx_rule, target_index) == 32 && offsetof(sx_rule, src_len) == 40,
                "sx_rule layout");
  struct Tables {
    int8_t bytes[2][64], cyc[64];
    Tables() {
      for (uint32_t op = 0; op < 64; op++) {
        for (int m = 0; m < 2; m++) bytes[m][op] = op < (uint32_t)OP_COUNT ? (int8_t)byte_size(op, m) : 0;
        cyc[op] = op < (uint32_t)OP_COUNT ? (int8_t)cycles(op) : 0;
      }
    }
  };
  static const Tables tab;
  const int8_t *tb = tab.bytes[model ? 1 : 0], *tc = tab.cyc;
  using v4u = uint32_t __attribute__((ext_vector_type(4)));
  auto rec = [&](uint64_t i, v4u &a, v4u &b, v4u &c) {
    const uint32_t *p = w + i * kPackedWords;
    const uint32_t w0 = p[0], w1 = p[1];
    const uint32_t s0 = inst_of14(w0), s1 = inst_of14(w0 >> 14), rp = inst_of14(w1);
    const uint32_t o0 = s0 & 63u, o1 = s1 & 63u, oc = rp & 63u;
    const int bs = tb[o0] + tb[o1] - tb[oc], cs = tc[o0] + tc[o1] - tc[oc];
    a = v4u{s0, s1, 0u, 0u};
    b = v4u{rp, 0u, 0u, 0u};
    c = v4u{p[2], 0u, 2u | 1u << 8 | ((w1 >> 14) & 0xFFu) << 16 | ((w0 >> 28) & 3u) << 24,
            (uint32_t)(uint16_t)bs | (uint32_t)(uint16_t)cs << 16};
  };
  v4u a, b, c;
  if ((reinterpret_cast<uintptr_t>(out) & 15) == 0) {
    v4u *o = reinterpret_cast<v4u *>(out);
    for (uint64_t i = 0; i < n; i++) {
      rec(i, a, b, c);
      __builtin_nontemporal_store(a, o + 3 * i);
      __builtin_nontemporal_store(b, o + 3 * i + 1);
      __builtin_nontemporal_store(c, o + 3 * i + 2);
    }
    std::atomic_thread_fence(std::memory_order_seq_cst);  // the streaming stores before any later use
    return;
  }
  for (uint64_t i = 0; i < n; i++) {
    rec(i, a, b, c);
    memcpy(reinterpret_cast<char *>(out + i), &a, 16);
    memcpy(reinterpret_cast<char *>(out + i) + 16, &b, 16);
    memcpy(reinterpret_cast<char *>(out + i) + 32, &c, 16);
  }
}

// The grouped sweep's sort of its jobs by group key: onesweep at every size
// above one block (up to 2^20 keys rocPRIM's default is a merge sort, about
// 20 dependent launches), Bits key bits per pass.
template <unsigned Bits>
using GroupSortCfg = rocprim::radix_sort_config<
    rocprim::default_config, rocprim::default_config,
    rocprim::radix_sort_onesweep_config<rocprim::kernel_config<1024, 16>, rocprim::kernel_config<1024, 16>, Bits,
                                        rocprim::block_radix_rank_algorithm::match>,
    0>;
hipError_t group_sort(int bits, void *tmp, size_t &bytes, const uint32_t *key, uint32_t *skey, const uint32_t *idx,
                      uint32_t *sidx, uint32_t n, hipStream_t s) {
  if (bits == 10)
    return rocprim::radix_sort_pairs<GroupSortCfg<10>>(tmp, bytes, key, skey, idx, sidx, n, 0, kGroupKeyBits, s);
  return rocprim::radix_sort_pairs<GroupSortCfg<8>>(tmp, bytes, key, skey, idx, sidx, n, 0, kGroupKeyBits, s);
}

// Exclusive scan of n uint32 counts into uint64 offsets; *total on device
// (in and out 16-byte aligned: the tiles move 16 bytes per access).
int scan_u32(sx_handle *h, const uint32_t *in, uint64_t n, uint64_t *out, uint64_t *d_total) {
  const uint64_t nb = (n + kScanTile - 1) / kScanTile;
  uint64_t *sums;
  ENSURE(B_SUMS, uint64_t, nb + 1, sums);
  if (n == 0) {
    HIPCHK(h, hipMemsetAsync(d_total, 0, sizeof(uint64_t), h->stream));
    return SX_OK;
  }
  if (((uintptr_t)in | (uintptr_t)out) & 15u) return set_err(h, SX_EINVAL, "scan buffers not 16-byte aligned");
  hipLaunchKernelGGL(k_tile_sums, dim3((unsigned)nb), dim3(256), 0, h->stream, in, n, sums, 0);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, h->stream, sums, nb, d_total);
  hipLaunchKernelGGL(k_tile_scan, dim3((unsigned)nb), dim3(256), 0, h->stream, in, n, (const uint64_t *)sums, out);
  HIPCHK(h, hipGetLastError());
  return SX_OK;
}

int ensure_l2(sx_handle *h) {
  if (h->d_l2) return SX_OK;
  const uint64_t raw = (uint64_t)kNumInsts * kNumInsts;
  const uint64_t nb = blocks_for(raw, 256);
  uint32_t *counts;
  uint64_t *offs, *total;
  ENSURE(B_GEN_COUNTS, uint32_t, nb, counts);
  ENSURE(B_GEN_OFFS, uint64_t, nb, offs);
  ENSURE(B_TOTAL, uint64_t, 1, total);
  hipLaunchKernelGGL(k_gen_count, dim3((unsigned)nb), dim3(256), 0, h->stream, counts);
  int rc = scan_u32(h, counts, nb, offs, total);
  if (rc) return rc;
  uint64_t n = 0;
  HIPCHK(h, hipMemcpyAsync(&n, total, sizeof(n), hipMemcpyDeviceToHost, h->stream));
  SYNC_WD("device sync");
  HIPCHK(h, hipMalloc(&h->d_l2, n * 2 * sizeof(uint32_t)));
  hipLaunchKernelGGL(k_gen_write, dim3((unsigned)nb), dim3(256), 0, h->stream, offs, h->d_l2);
  HIPCHK(h, hipGetLastError());
  SYNC_WD("device sync");
  h->n_l2 = n;
  return SX_OK;
}


// One QuickCheck pass over targets [0, nt) x candidate slots [0, nc) with a
// given launch configuration (blocks_per_cu 0: grid form with `tile` targets
// per block row; else persistent form, num_cu x blocks_per_cu blocks).
int launch_qc(sx_handle *h, uint32_t tile, uint32_t blocks_per_cu, const uint32_t *tv0, const uint2 *tfp, uint64_t nt,
              const uint2 *cfp, uint64_t nc, uint32_t pm_masked, uint2 *surv, uint32_t surv_cap,
              uint32_t *counters) {
  if (blocks_per_cu == 0 && tile == 0) {
    if (!h->qj_dir) return set_err(h, SX_EINVAL, "QuickCheck join without its candidate table");
    hipLaunchKernelGGL(k_qc_join, dim3(blocks_for(nt, 256)), dim3(256), 0, h->stream, tv0, tfp, (uint32_t)nt,
                       h->qj_hash, h->qj_slot, h->qj_dir, h->qj_dk, cfp, pm_masked, surv, surv_cap / kQcRegions,
                       counters);
  } else if (blocks_per_cu == 0) {
    dim3 grid(blocks_for(nc, 256), (unsigned)((nt + tile - 1) / tile));
    hipLaunchKernelGGL(k_quickcheck, grid, dim3(256), 0, h->stream, tv0, tfp, (uint32_t)nt, tile, cfp,
                       (uint32_t)nc, pm_masked, surv, surv_cap / kQcRegions, counters);
  } else {
    uint32_t *heads;
    ENSURE(B_HEADS, uint32_t, 8, heads);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, h->stream, heads, (uint64_t)8, 0u);
    const uint32_t ncb = blocks_for(nc, 256), ntiles = (uint32_t)((nt + tile - 1) / tile);
    hipLaunchKernelGGL(k_qc_persistent, dim3((unsigned)(h->num_cu * blocks_per_cu)), dim3(256), 0, h->stream, tv0,
                       tfp, (uint32_t)nt, tile, cfp, (uint32_t)nc, pm_masked, surv, surv_cap / kQcRegions, counters,
                       heads, ncb, ntiles);
  }
  HIPCHK(h, hipGetLastError());
  return SX_OK;
}

// The QuickCheck autotuner: on the first segment of a (target_len,
// max_cand_len) shape, time every candidate configuration on a slice of the
// real work (about 4e9 pairs) and keep the fastest for the handle.  The
// survivors of these runs are discarded; the real pass runs afterwards.
int tune_qc(sx_handle *h, int tl, int ml, const uint32_t *tv0, const uint2 *tfp, uint64_t nt, const uint2 *cfp,
            uint64_t nc, uint32_t pm_masked, uint2 *surv, uint32_t surv_cap, uint32_t *counters,
            sx_handle::QcChoice *out) {
  const int big = nt >= 65536;
  for (const auto &c : h->qc_tuned)
    if (c.tl == tl && c.ml == ml && c.big == big) {
      *out = c;
      return SX_OK;
    }
  const uint64_t nc_t = std::min<uint64_t>(nc, 1u << 18);
  const uint64_t nt_t = std::min<uint64_t>(nt, std::max<uint64_t>(16384, 4000000000ull / nc_t));
  struct Cand {
    uint32_t tile, bpc;
  };
  // {0, 0}: the hash join (over all nc candidates; the others take the first
  // nc_t), compared per pair
  const Cand cands[] = {{0, 0},    {1024, 0}, {2048, 0}, {4096, 0}, {8192, 0}, {2048, 2},
                        {2048, 4}, {2048, 8}, {4096, 2}, {4096, 4}, {4096, 8}, {8192, 4}};
  float best = 1e30f;
  sx_handle::QcChoice choice{tl, ml, big, 2048, 0};
  for (int rep = 0; rep < 2; rep++)  // the first sweep warms up code and caches
    for (const Cand &c : cands) {
      const bool join = c.tile == 0 && c.bpc == 0;
      if (join && !h->qj_dir) continue;
      HIPCHK(h, hipEventRecord(h->kev[0], h->stream));
      int rc = launch_qc(h, c.tile, c.bpc, tv0, tfp, nt_t, cfp, join ? nc : nc_t, pm_masked, surv, surv_cap,
                         counters);
      if (rc) return rc;
      HIPCHK(h, hipEventRecord(h->kev[1], h->stream));
      rc = sync_wd(h, "QuickCheck autotune");
      if (rc) return rc;
      float ms = 0;
      HIPCHK(h, hipEventElapsedTime(&ms, h->kev[0], h->kev[1]));
      if (join) ms *= (float)nc_t / (float)nc;  // per pair, like the slices of the others
      trace(h, "  autotune QuickCheck tile %u blocks/CU %u: %.3f ms (per %llu candidates)", c.tile, c.bpc, ms,
            (unsigned long long)nc_t);
      if (rep == 1 && ms < best) {
        best = ms;
        choice.tile = c.tile;
        choice.blocks_per_cu = c.bpc;
      }
    }
  h->qc_tuned.push_back(choice);
  *out = choice;
  return SX_OK;
}

// One launch of the sweep over chunks [c0, c0 + n) (a wave per chunk) in the
// handle's interpreter variant.  chunk_job may be null (the waves then find
// their job by searching the chunk offsets).
int launch_exh_run(sx_handle *h, const Job *jobs, uint32_t nj, const uint64_t *offs, const uint32_t *chunk_job,
                   uint64_t c0, uint64_t n, uint32_t *fail, unsigned long long *evald) {
  const dim3 grid((unsigned)((n + 3) / 4));
  if (h->exh_pk) {
    if (h->exh_k == 1)
      hipLaunchKernelGGL((k_exh_run<1, true>), grid, dim3(256), 0, h->stream, jobs, nj, offs, chunk_job, c0, n,
                         fail, evald);
    else
      hipLaunchKernelGGL((k_exh_run<2, true>), grid, dim3(256), 0, h->stream, jobs, nj, offs, chunk_job, c0, n,
                         fail, evald);
  } else {
    if (h->exh_k == 1)
      hipLaunchKernelGGL((k_exh_run<1, false>), grid, dim3(256), 0, h->stream, jobs, nj, offs, chunk_job, c0, n,
                         fail, evald);
    else
      hipLaunchKernelGGL((k_exh_run<2, false>), grid, dim3(256), 0, h->stream, jobs, nj, offs, chunk_job, c0, n,
                         fail, evald);
  }
  HIPCHK(h, hipGetLastError());
  return SX_OK;
}

// Run the exhaustive stage on nj jobs already in B_JOBS; the verdicts are in
// B_FAIL (0 = pass).  *chunks_out gets the number of chunks (statistics).
// planned: the jobs' plans, chunk counts and fail flags are already written
// (by the kernel that made the jobs).
// ptab / pm_masked: the plan table of the call (or null) and its relaxed mask.
// keyed: the kernel that planned the jobs also wrote their group keys into
// B_GKEY / B_GIDX (with the grouped sweep on).
template <int K>
void launch_exh_group_k(bool ent, dim3 gg, hipStream_t s, const Job *jobs, const uint4 *jent, const PlanEntry *ptab,
                        uint32_t pm_masked, const uint32_t *skey, const uint32_t *sidx, uint32_t nj, uint32_t *fail,
                        unsigned long long *gev) {
  if (ent)
    hipLaunchKernelGGL((k_exh_group<K, true>), gg, dim3(256), 0, s, jobs, jent, ptab, pm_masked, skey, sidx, nj, fail,
                       gev);
  else
    hipLaunchKernelGGL((k_exh_group<K, false>), gg, dim3(256), 0, s, jobs, jent, ptab, pm_masked, skey, sidx, nj,
                       fail, gev);
}
void launch_exh_group(int k, bool ent, dim3 gg, hipStream_t s, const Job *jobs, const uint4 *jent,
                      const PlanEntry *ptab, uint32_t pm_masked, const uint32_t *skey, const uint32_t *sidx,
                      uint32_t nj, uint32_t *fail, unsigned long long *gev) {
  if (k == 8)
    launch_exh_group_k<8>(ent, gg, s, jobs, jent, ptab, pm_masked, skey, sidx, nj, fail, gev);
  else if (k == 4)
    launch_exh_group_k<4>(ent, gg, s, jobs, jent, ptab, pm_masked, skey, sidx, nj, fail, gev);
  else if (k == 2)
    launch_exh_group_k<2>(ent, gg, s, jobs, jent, ptab, pm_masked, skey, sidx, nj, fail, gev);
  else
    launch_exh_group_k<1>(ent, gg, s, jobs, jent, ptab, pm_masked, skey, sidx, nj, fail, gev);
}

// jent (with ptab): the keyed jobs are compact records there (k_bulk_jobs);
// B_JOBS holds only the others.
int run_exhaustive(sx_handle *h, uint32_t nj, int mode, uint64_t *chunks_out, bool planned = false,
                   const PlanEntry *ptab = nullptr, uint32_t pm_masked = 0, bool keyed = false,
                   const uint4 *jent = nullptr) {
  Job *jobs = (Job *)h->bufs[B_JOBS].p;
  uint32_t *chunks, *fail;
  uint64_t *offs, *total;
  ENSURE(B_CHUNKS, uint32_t, nj + 1, chunks);
  ENSURE(B_OFFS, uint64_t, nj + 1, offs);
  ENSURE(B_FAIL, uint32_t, nj + 1, fail);
  ENSURE(B_TOTAL, uint64_t, 1, total);
  unsigned long long *evald;  // evaluated states, added up across calls (three banks)
  ENSURE(B_CURSOR, unsigned long long, 3 * kEvalShards, evald);
  uint32_t *key = nullptr, *idx = nullptr;  // group keys of the jobs (grouped sweep)
  if (h->exh_group && nj) {
    ENSURE(B_GKEY, uint32_t, nj, key);
    ENSURE(B_GIDX, uint32_t, nj, idx);
  }
  if (!planned) {
    hipLaunchKernelGGL(k_exh_plan, dim3(blocks_for(nj, 256)), dim3(256), 0, h->stream, jobs, nj, mode, ptab, pm_masked,
                       chunks, fail, key, idx);
    keyed = key != nullptr;
  }
  HIPCHK(h, hipGetLastError());
  if (h->exh_group && nj) {
    // jobs with small plans: one job per lane, sorted by (opcodes, mask); they
    // leave the state-parallel sweep below with zero chunks
    uint32_t *skey, *sidx;
    ENSURE(B_GKEY_S, uint32_t, nj, skey);
    ENSURE(B_GIDX_S, uint32_t, nj, sidx);
    if (!keyed)
      hipLaunchKernelGGL(k_group_keys, dim3(blocks_for(nj, 256)), dim3(256), 0, h->stream, jobs, nj, chunks, key, idx);
    HIPCHK(h, hipGetLastError());
    size_t tmp_bytes = 0;
    HIPCHK(h, group_sort(h->sort_bits, nullptr, tmp_bytes, key, skey, idx, sidx, nj, h->stream));
    uint8_t *tmp;
    ENSURE(B_GSORT, uint8_t, tmp_bytes + 1, tmp);
    HIPCHK(h, group_sort(h->sort_bits, tmp, tmp_bytes, key, skey, idx, sidx, nj, h->stream));
    size_t ts = 0;
    int rc = timed_begin(h, 2, &ts);
    if (rc) return rc;
    const dim3 gg(blocks_for(nj, kWave));
    unsigned long long *gev = evald + kEvalShards;
    launch_exh_group(h->exh_group, jent != nullptr, gg, h->stream, jobs, jent, ptab, pm_masked, skey, sidx, nj, fail,
                     gev);
    HIPCHK(h, hipGetLastError());
    rc = timed_end(h, ts);
    if (rc) return rc;
  }
  int rc = scan_u32(h, chunks, nj, offs, total);
  if (rc) return rc;
  uint64_t tot = 0;
  rc = read_back(h, {{&tot, total, sizeof(tot)}}, "exhaustive plan");
  if (rc) return rc;
  *chunks_out = tot;
  trace(h, "  exhaustive: %u jobs, %llu chunks", nj, (unsigned long long)tot);
  // a direct chunk -> job map while it stays small (huge sweeps use the search)
  uint32_t *chunk_job = nullptr;
  if (tot <= (1ull << 26)) {
    ENSURE(B_CHUNKMAP, uint32_t, tot + 1, chunk_job);
    if (tot <= 2ull * nj)
      hipLaunchKernelGGL(k_chunk_map, dim3(blocks_for(nj, 256)), dim3(256), 0, h->stream, offs, chunks, nj, chunk_job);
    else
      hipLaunchKernelGGL(k_chunk_map_search, dim3((unsigned)std::min<uint64_t>(blocks_for(tot, 256), 1u << 16)),
                         dim3(256), 0, h->stream, offs, chunks, nj, tot, chunk_job);
    HIPCHK(h, hipGetLastError());
  }
  const uint64_t slice = 1ull << 20;  // chunks per launch (2^34 states)
  for (uint64_t c0 = 0; c0 < tot; c0 += slice) {
    uint64_t n = std::min<uint64_t>(slice, tot - c0);
    size_t ts = 0;
    rc = timed_begin(h, 1, &ts);
    if (rc) return rc;
    rc = launch_exh_run(h, jobs, nj, offs, chunk_job, c0, n, fail, evald);
    if (rc) return rc;
    rc = timed_end(h, ts);
    if (rc) return rc;
  }
  return SX_OK;
}

// The length-3 stage of one segment (first hit): the entries still open after
// the candidates of length 1 and 2 against all length-3 candidates, in
// sweeps of up to kL3MaxTargets open targets; resolved entries get their rule
// and flag in the segment's slots.
int run_l3(sx_handle *h, uint32_t nt, const uint32_t *tw, uint32_t tl, const uint2 *tfp, const uint8_t *tbytes,
           const uint8_t *tcyc, uint32_t *rule_flag, sx_rule *rules, uint32_t dead, int model, int exh_mode,
           uint64_t tbase, sx_stats *st) {
  const uint32_t pm_masked = kPMask & ~dead, mlo = 0x00FFFFFFu | (pm_masked << 24);
  uint32_t *list, *hk, *cnt;
  unsigned long long *best;
  ENSURE(B_L3_LIST, uint32_t, nt, list);
  ENSURE(B_L3_HK, uint32_t, nt, hk);
  ENSURE(B_L3_CNT, uint32_t, 4, cnt);
  ENSURE(B_L3_BEST, unsigned long long, 2 * (uint64_t)nt, best);
  HIPCHK(h, hipMemsetAsync(cnt, 0, 16, h->stream));
  HIPCHK(h, hipMemsetAsync(best, 0xFF, 2 * (uint64_t)nt * 8, h->stream));
  hipLaunchKernelGGL(k_l3_open, dim3(blocks_for(nt, 256)), dim3(256), 0, h->stream, rule_flag, nt, dead ? 1 : 0,
                     tbytes, tfp, mlo, list, hk, cnt);
  HIPCHK(h, hipGetLastError());
  uint32_t c[2] = {0, 0};
  HIPCHK(h, hipMemcpyAsync(c, cnt, 8, hipMemcpyDeviceToHost, h->stream));
  SYNC_WD("device sync");
  const uint32_t nopen = c[0];
  const int max_tbytes = (int)c[1];
  trace(h, "  length-3 candidates: %u open targets (at most %d bytes)", nopen, max_tbytes);
  uint32_t raw_cap = 1u << 20;
  for (uint32_t b0 = 0; b0 < nopen; b0 += kL3MaxTargets) {
    const uint32_t n = std::min(kL3MaxTargets, nopen - b0);
    uint32_t lg = 6;
    while ((1u << lg) < 4 * n) lg++;
    const uint32_t mask = (1u << lg) - 1u;
    uint32_t *keys, *bloom;
    int32_t *head, *next;
    ENSURE(B_L3_KEYS, uint32_t, mask + 1, keys);
    ENSURE(B_L3_HEAD, int32_t, mask + 1, head);
    ENSURE(B_L3_NEXT, int32_t, n, next);
    ENSURE(B_L3_BLOOM, uint32_t, kL3BloomWords, bloom);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(mask + 1, 256)), dim3(256), 0, h->stream, keys, mask + 1, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(mask + 1, 256)), dim3(256), 0, h->stream, (uint32_t *)head,
                       mask + 1, 0xFFFFFFFFu);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(kL3BloomWords, 256)), dim3(256), 0, h->stream, bloom,
                       kL3BloomWords, 0u);
    hipLaunchKernelGGL(k_l3_insert, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, hk + b0, n, mask, keys, head,
                       next, bloom);
    HIPCHK(h, hipGetLastError());
    // the sweep, in launches of 2^20 prefixes (2^17 waves); repeated with a
    // larger hit buffer if it overflowed
    uint2 *raw;
    uint32_t nraw = 0;
    for (;;) {
      ENSURE(B_L3_RAW, uint2, raw_cap, raw);
      HIPCHK(h, hipMemsetAsync(cnt + 2, 0, 4, h->stream));
      const uint64_t per = 1u << 20;
      for (uint64_t p0 = 0; p0 < kL3Prefixes; p0 += per) {
        const uint64_t np = std::min<uint64_t>(per, kL3Prefixes - p0);
        hipLaunchKernelGGL(k_l3_join, dim3((unsigned)((np + 31) / 32)), dim3(256), 0, h->stream, p0, np, keys, mask,
                           bloom, max_tbytes, model, mlo, raw, raw_cap, cnt + 2);
        HIPCHK(h, hipGetLastError());
      }
      HIPCHK(h, hipMemcpyAsync(&nraw, cnt + 2, 4, hipMemcpyDeviceToHost, h->stream));
      SYNC_WD("length-3 sweep");
      if (nraw <= raw_cap) break;
      if (nraw > (1u << 30)) return set_err(h, SX_ENOMEM, "length-3 hits (%u)", nraw);
      raw_cap = nraw + nraw / 4;
    }
    st->qc_pairs += (uint64_t)n * kL3Prefixes * kNumInsts;
    st->qc_raw_hits += nraw;
    trace(h, "  length-3 sweep of %u targets: %u hits", n, nraw);
    // hits -> jobs in chunks, the sweep, the smallest passing index per entry
    const uint32_t chunk = 1u << 18;
    for (uint32_t r0 = 0; r0 < nraw; r0 += chunk) {
      const uint32_t nr = std::min(chunk, nraw - r0);
      uint32_t job_cap = 4 * nr, nj = 0;
      for (;;) {
        Job *jobs;
        uint64_t *c3;
        ENSURE(B_JOBS, Job, job_cap, jobs);
        ENSURE(B_L3_C3, uint64_t, job_cap, c3);
        HIPCHK(h, hipMemsetAsync(cnt + 3, 0, 4, h->stream));
        hipLaunchKernelGGL(k_l3_expand, dim3(blocks_for(nr, 256)), dim3(256), 0, h->stream, raw + r0, nr, head, next,
                           list + b0, tw, tl, tfp, tbytes, rule_flag, dead ? 1 : 0, pm_masked, model, jobs, c3,
                           job_cap, cnt + 3);
        HIPCHK(h, hipGetLastError());
        HIPCHK(h, hipMemcpyAsync(&nj, cnt + 3, 4, hipMemcpyDeviceToHost, h->stream));
        SYNC_WD("device sync");
        if (nj <= job_cap) break;
        job_cap = nj;
      }
      if (nj == 0) continue;
      st->qc_hits += nj;
      st->exh_pairs += nj;
      st->exh_rounds++;
      uint64_t chunks = 0;
      int rc = run_exhaustive(h, nj, exh_mode, &chunks, false);
      if (rc) return rc;
      hipLaunchKernelGGL(k_l3_best, dim3(blocks_for(nj, 256)), dim3(256), 0, h->stream, (const Job *)h->bufs[B_JOBS].p,
                         nj, (const uint32_t *)h->bufs[B_FAIL].p, (const uint64_t *)h->bufs[B_L3_C3].p, tfp, best);
      HIPCHK(h, hipGetLastError());
    }
  }
  hipLaunchKernelGGL(k_l3_emit, dim3(blocks_for(2 * (uint64_t)nt, 256)), dim3(256), 0, h->stream, best, nt, tw, tl,
                     tfp, tbytes, tcyc, model, tbase, rules, rule_flag);
  HIPCHK(h, hipGetLastError());
  return SX_OK;
}

// Sums of the three banks of evaluated-state shards: out[b] = sum of bank b.
static_assert(kEvalShards == 1024, "one shard per thread of one block");
__global__ __launch_bounds__(1024) void k_sum_shards(const unsigned long long *ev, unsigned long long *out) {
  __shared__ unsigned long long s[3][16];
  const uint32_t t = threadIdx.x, w = t >> 6;
  for (int b = 0; b < 3; b++) {
    const uint64_t x = wave_sum64(ev[b * kEvalShards + t]);
    if (lane_id() == 0) s[b][w] = x;
  }
  __syncthreads();
  if (t < 3) {
    uint64_t x = 0;
    for (int k = 0; k < 16; k++) x += s[t][k];
    out[t] = x;
  }
}

// The evaluated-state counters (sharded, three banks: the state-parallel
// sweep's states, the grouped sweep's states, the grouped sweep's jobs):
// cleared, and summed on the host.
int clear_evaluated(sx_handle *h) {
  unsigned long long *evald;
  ENSURE(B_CURSOR, unsigned long long, 3 * kEvalShards, evald);
  HIPCHK(h, hipMemsetAsync(evald, 0, 3 * kEvalShards * sizeof(unsigned long long), h->stream));
  return SX_OK;
}
// The three banks summed on the device first: a read-back of a few bytes is
// a kernel copy on the library's stream, while one of some kilobytes goes to
// the copy engine, behind the rules' copy to the host.  bad / bad_d: a 32-bit
// device flag read in the same wait.
int read_evaluated(sx_handle *h, uint64_t *sum, uint64_t *sum_group, uint64_t *group_jobs, uint32_t *bad,
                   const uint32_t *bad_d, PendingRules *pend) {
  unsigned long long *evald, *sums;
  ENSURE(B_CURSOR, unsigned long long, 3 * kEvalShards, evald);
  ENSURE(B_EVSUM, unsigned long long, 3, sums);
  hipLaunchKernelGGL(k_sum_shards, dim3(1), dim3(kEvalShards), 0, h->stream, evald, sums);
  HIPCHK(h, hipGetLastError());
  unsigned long long tot[3] = {0, 0, 0};
  int rc = bad ? read_back(h, {{tot, sums, sizeof(tot)}, {bad, bad_d, 4}}, "exhaustive sweep", pend)
               : read_back(h, {{tot, sums, sizeof(tot)}}, "exhaustive sweep", pend);
  if (rc) return rc;
  *sum = tot[0] + tot[1];
  *sum_group = tot[1];
  *group_jobs = tot[2];
  return SX_OK;
}

}  // namespace

// ------------------------------------------------------------------- ABI --

extern "C" int sx_abi_version(void) { return SX_ABI_VERSION; }

namespace {
// 16 independent chains per lane, 32 v_xor_b32 per iteration (2 per chain):
// the int32 VALU rate with nothing else in the way but the loop branch.
__global__ __launch_bounds__(256) void k_int32_rate(uint32_t *out, uint32_t iters, uint32_t k) {
  uint32_t a[16];
#pragma unroll
  for (int j = 0; j < 16; j++) a[j] = threadIdx.x * 16 + j;
  for (uint32_t i = 0; i < iters; i++) {
#pragma unroll
    for (int j = 0; j < 16; j++) a[j] ^= k + j;
#pragma unroll
    for (int j = 0; j < 16; j++) a[j] ^= a[(j + 1) & 15];
  }
  uint32_t s = 0;
#pragma unroll
  for (int j = 0; j < 16; j++) s ^= a[j];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}
}  // namespace

extern "C" int sx_measure_int32_rate(sx_handle *h, double *ops_per_s) {
  if (!h || !ops_per_s) return SX_EINVAL;
  if (h->poisoned) return set_err(h, SX_EHIP, "handle unusable after a device timeout");
  HIPCHK(h, hipSetDevice(h->device));
  const unsigned blocks = (unsigned)h->num_cu * 32;
  const uint32_t iters = 5000;
  uint32_t *out;
  ENSURE(B_SCRATCH, uint32_t, (size_t)blocks * 256, out);
  double best = 0;
  for (int rep = 0; rep < 4; rep++) {  // the first run warms up
    HIPCHK(h, hipEventRecord(h->kev[0], h->stream));
    hipLaunchKernelGGL(k_int32_rate, dim3(blocks), dim3(256), 0, h->stream, out, iters, 7u);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipEventRecord(h->kev[1], h->stream));
    int rc = sync_wd(h, "int32 rate");
    if (rc) return rc;
    float ms = 0;
    HIPCHK(h, hipEventElapsedTime(&ms, h->kev[0], h->kev[1]));
    const double r = (double)blocks * 256 * iters * 32 / (ms * 1e-3);
    if (rep > 0 && r > best) best = r;
  }
  *ops_per_s = best;
  return SX_OK;
}

// ---- sx_check_pairs_noalloc: the pair check without device allocations.
// Its buffers live in `scratch`, a file-scope device array of fixed size
// that every GPU the library runs on holds until the process ends.  Host
// code never passes `scratch` itself to a kernel or a copy (in host code the
// name is a host-side stand-in, not a device address): sx_open looks up the
// array's address on its device with hipGetSymbolAddress, once per device,
// and the entry point uses that pointer.  All calls on one GPU share the
// array, so calls that overlap on one GPU race.
constexpr uint64_t kNaMaxPairs = SX_NOALLOC_MAX_PAIRS;  // pairs per call
constexpr size_t na_align(size_t x) { return (x + 255) & ~(size_t)255; }
struct NaLayout {  // byte offsets into the scratch array
  size_t tw, cw, tl, cl, qc, mid, out, jobs, chunks, offs, fail, sums, total, evald, bytes;
};
constexpr NaLayout na_layout() {
  const size_t n = kNaMaxPairs;
  NaLayout L{};
  size_t o = 0;
  L.tw = o;     o += na_align(16 * n);
  L.cw = o;     o += na_align(16 * n);
  L.tl = o;     o += na_align(n);
  L.cl = o;     o += na_align(n);
  L.qc = o;     o += na_align(n);
  L.mid = o;    o += na_align(n);
  L.out = o;    o += na_align(n);
  L.jobs = o;   o += na_align(sizeof(Job) * n);
  L.chunks = o; o += na_align(4 * (n + 1));
  L.offs = o;   o += na_align(8 * (n + 1));
  L.fail = o;   o += na_align(4 * (n + 1));
  L.sums = o;   o += na_align(8 * (n / kScanTile + 2));
  L.total = o;  o += na_align(8);
  L.evald = o;  o += na_align(8 * kEvalShards);
  L.bytes = o;
  return L;
}
constexpr NaLayout kNa = na_layout();
static_assert(kNa.bytes <= (12u << 20), "the no-alloc scratch stays at 12 MiB");
static __device__ __align__(256) uint8_t scratch[kNa.bytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices];      // per device: the device address of `scratch`
hipError_t g_scratch_rc[kMaxDevices];  // the lookup's status (valid once g_scratch_ptr or g_scratch_tried is set)
bool g_scratch_tried[kMaxDevices];

// With `device` current: look up the address of `scratch` on it (once per
// device; later handles on the device reuse the result).
hipError_t scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_tried[device]) {
    void *p = nullptr;
    g_scratch_rc[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (g_scratch_rc[device] != hipSuccess) (void)hipGetLastError();
    else g_scratch_ptr[device] = p;
    g_scratch_tried[device] = true;
  }
  return g_scratch_rc[device];
}
}  // namespace

extern "C" int sx_open(int device, sx_handle **out) {
  if (!out) return SX_EINVAL;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) {
    (void)hipGetLastError();
    return SX_ENODEV;
  }
  sx_handle *h = new sx_handle();
  h->device = device;
  if (hipSetDevice(device) != hipSuccess) {
    delete h;
    return SX_ENODEV;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) h->num_cu = prop.multiProcessorCount;
  h->err.reserve(512);  // set_err's messages then fit without a heap allocation
  h->scratch_rc = scratch_init(device);
  h->scratch = h->scratch_rc == hipSuccess ? static_cast<uint8_t *>(g_scratch_ptr[device]) : nullptr;
  if (const char *e = getenv("SX_SORT_BITS")) h->sort_bits = atoi(e) == 10 ? 10 : 8;
  if (const char *e = getenv("SX_BULK_AFTER")) h->bulk_after = std::max(-2, atoi(e));
  if (const char *e = getenv("SX_EXH_K")) h->exh_k = atoi(e) == 1 ? 1 : 2;
  if (const char *e = getenv("SX_EXH_PK")) h->exh_pk = atoi(e) != 0;
  if (const char *e = getenv("SX_EXH_GROUP")) h->exh_group = atoi(e);
  if (const char *e = getenv("SX_QCM_JOIN")) h->qcm_join = atoi(e) != 0;
  if (const char *e = getenv("SX_FP_SPLIT")) h->fp_split = atoi(e) != 0;
  if (const char *e = getenv("SX_COPY_STREAMS")) h->copy_streams = atoi(e) == 1 ? 1 : 2;
  if (const char *e = getenv("SX_FETCH_THREADS")) h->fetch_threads = std::min(64, std::max(1, atoi(e)));
  if (const char *e = getenv("SX_SEG_LOG")) h->seg_log = std::min(24, std::max(16, atoi(e)));
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&h->copy_stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&h->copy_stream2, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&h->kev[0]) != hipSuccess || hipEventCreate(&h->kev[1]) != hipSuccess ||
      hipEventCreateWithFlags(&h->seg_ev, hipEventDisableTiming) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void **>(&h->pin), kPinBytes, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    sx_close(h);
    return SX_EHIP;
  }
  *out = h;
  return SX_OK;
}

extern "C" int sx_close(sx_handle *h) {
  if (!h) return SX_EINVAL;
  {
    std::lock_guard<std::mutex> lk(h->fetch_pool.mu);
    h->fetch_pool.stop = true;
    h->fetch_pool.cv.notify_all();
  }
  for (auto &t : h->fetch_pool.th) t.join();
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->copy_stream) (void)hipStreamSynchronize(h->copy_stream);
  if (h->copy_stream2) (void)hipStreamSynchronize(h->copy_stream2);
  for (auto &b : h->bufs)
    if (b.p) (void)hipFree(b.p);
  if (h->d_l2) (void)hipFree(h->d_l2);
  if (h->d_rules) (void)hipFree(h->d_rules);
  if (h->h_rules) (void)hipHostFree(h->h_rules);
  if (h->pin) (void)hipHostFree(h->pin);
  for (hipEvent_t e : h->kev)
    if (e) (void)hipEventDestroy(e);
  for (auto &t : h->timed) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  if (h->seg_ev) (void)hipEventDestroy(h->seg_ev);
  for (hipEvent_t e : h->piece_ev) (void)hipEventDestroy(e);
  if (h->copy_stream) (void)hipStreamDestroy(h->copy_stream);
  if (h->copy_stream2) (void)hipStreamDestroy(h->copy_stream2);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return SX_OK;
}

extern "C" int sx_last_error(sx_handle *h, char *buf, size_t cap) {
  if (!h || !buf || cap == 0) return SX_EINVAL;
  snprintf(buf, cap, "%s", h->err.c_str());
  return SX_OK;
}

extern "C" int sx_pruned_targets(sx_handle *h, uint64_t first, uint64_t n, uint32_t *out_words,
                                 uint64_t *total) {
  if (!h) return SX_EINVAL;
  HIPCHK(h, hipSetDevice(h->device));
  int rc = ensure_l2(h);
  if (rc) return rc;
  if (total) *total = h->n_l2;
  if (!out_words || n == 0) return SX_OK;
  if (first > h->n_l2 || n > h->n_l2 - first) return set_err(h, SX_EINVAL, "range out of bounds");
  HIPCHK(h, hipMemcpy(out_words, h->d_l2 + 2 * first, n * 2 * sizeof(uint32_t), hipMemcpyDeviceToHost));
  return SX_OK;
}

static int check_words(sx_handle *h, const uint32_t *w, uint64_t n) {
  for (uint64_t i = 0; i < n; i++)
    if ((w[i] & 0xFFFFu) >= OP_COUNT) return set_err(h, SX_EOPCODE, "opcode %u out of range", w[i] & 0xFFFFu);
  return SX_OK;
}

extern "C" int sx_enumerate(sx_handle *h, const sx_enum_cfg *cfg_in, sx_rule *out, uint64_t cap,
                            uint64_t *n_out, sx_stats *stats) {
  if (!h || !cfg_in) return SX_EINVAL;
  if (h->poisoned) return set_err(h, SX_EHIP, "handle unusable after a device timeout");
  if (cfg_in->struct_size != sizeof(sx_enum_cfg) && cfg_in->struct_size != SX_ENUM_CFG_SIZE_V1)
    return set_err(h, SX_EINVAL, "sx_enum_cfg size mismatch");
  // a caller of the first layout leaves the later fields at 0
  sx_enum_cfg cfg_local;
  memset(&cfg_local, 0, sizeof(cfg_local));
  memcpy(&cfg_local, cfg_in, cfg_in->struct_size);
  const sx_enum_cfg *cfg = &cfg_local;
  // search.go:26-50: --max-target L runs every target length 2..L with
  // candidates of length 1..L-1.  Candidates of length 1 and 2 are slots of
  // the pool (ml); length 3 (ml_req 3) is a stage of its own after them.
  // Length-4 targets come as a list.
  const int tl = cfg->target_len, ml_req = cfg->max_cand_len, ml = std::min(ml_req, 2);
  if (tl < 2 || tl > 4 || ml_req < 1 || ml_req > 3)
    return set_err(h, SX_EINVAL, "target_len %d / max_cand_len %d unsupported", tl, ml_req);
  if (tl == 4 && !cfg->targets) return set_err(h, SX_EINVAL, "length-4 targets must be given as a list");
  if (ml_req == 3 && cfg->hits) return set_err(h, SX_EINVAL, "length-3 candidates need first-hit mode");
  if (cfg->bytes_model < 0 || cfg->bytes_model > 1 || cfg->hits < 0 || cfg->hits > 1 ||
      cfg->exh_mode < 0 || cfg->exh_mode > 1 || cfg->dead_flags > 0xFF)
    return set_err(h, SX_EINVAL, "bad mode");
  // implicit targets: length 2 = the cached pruned list, indexed by position;
  // length 3 = ranges of raw Go-order triple indices, pruned on the device
  const bool implicit3 = !cfg->targets && tl == 3;
  const uint64_t raw3 = (uint64_t)kNumInsts * kNumInsts * kNumInsts;
  HIPCHK(h, hipSetDevice(h->device));
  auto wall0 = std::chrono::steady_clock::now();
  const double qc_ms0 = h->k_qc_ms, exh_ms0 = h->k_exh_ms, grp_ms0 = h->k_grp_ms;
  const uint64_t qc_l0 = h->k_qc_launches, exh_l0 = h->k_exh_launches, grp_l0 = h->k_grp_launches;
  sx_stats st;
  memset(&st, 0, sizeof(st));
  h->rules.clear();
  h->empty_rules.clear();
  HIPCHK(h, hipStreamSynchronize(h->copy_stream));  // the last call's rule copies read d_rules
  HIPCHK(h, hipStreamSynchronize(h->copy_stream2));
  h->d_rules_n = 0;
  h->pieces.clear();
  h->rules_on_device = !cfg->hits;
  h->rec_words = 12;
  {
    int rc = clear_evaluated(h);
    if (rc) return rc;
  }

  // ---- targets
  const uint32_t *d_tsrc = nullptr;
  uint64_t nt_all = 0, tbase = 0;
  if (cfg->targets) {
    int rc = check_words(h, cfg->targets, cfg->n_targets * tl);
    if (rc) return rc;
    nt_all = cfg->n_targets;
  } else if (implicit3) {
    uint64_t a = cfg->first_target, b = cfg->end_target;
    if (a == 0 && b == 0) b = raw3;
    if (a > b || b > raw3) return set_err(h, SX_EINVAL, "target range out of bounds");
    tbase = a;
    nt_all = b - a;  // raw indices; each segment keeps its unpruned triples
  } else {
    int rc = ensure_l2(h);
    if (rc) return rc;
    uint64_t a = cfg->first_target, b = cfg->end_target;
    if (a == 0 && b == 0) b = h->n_l2;
    if (a > b || b > h->n_l2) return set_err(h, SX_EINVAL, "target range out of bounds");
    tbase = a;
    nt_all = b - a;
    d_tsrc = h->d_l2 + 2 * a;
  }
  if (h->rules_on_device && tl == 2 && ml == 1 && tbase + nt_all <= 0xFFFFFFFFull) h->rec_words = kPackedWords;
  h->fetch_model = cfg->bytes_model;

  // ---- candidates: slots in Go order (rank = slot)
  const uint64_t nc = ml == 1 ? (uint64_t)kNumInsts : (uint64_t)kNumInsts + (uint64_t)kNumInsts * kNumInsts;
  uint32_t *cwords, *rank, *badop;
  uint2 *cfp, *cmh;
  uint8_t *cbytes, *ccyc, *cprune;
  const uint32_t pm_masked = kPMask & ~cfg->dead_flags;
  const uint32_t lo_mask = 0x00FFFFFFu | (pm_masked << 24);
  ENSURE(B_CWORDS, uint32_t, 2 * nc, cwords);
  ENSURE(B_RANK, uint32_t, nc, rank);
  ENSURE(B_CFP, uint2, nc * kFpVecs, cfp);
  ENSURE(B_CMH, uint2, nc, cmh);
  ENSURE(B_CBYTES, uint8_t, nc, cbytes);
  ENSURE(B_CCYC, uint8_t, nc, ccyc);
  ENSURE(B_CPRUNE, uint8_t, nc, cprune);
  ENSURE(B_BADOP, uint32_t, 1, badop);
  // stage events: two sets for the segments, used in turn, so that a
  // segment's stage times are read once a later wait has passed them (no
  // wait of their own)
  hipEvent_t e0, es, seg_evs[2][4];
  bool seg_evs_pending[2] = {false, false};
  HIPCHK(h, hipEventCreate(&e0));
  HIPCHK(h, hipEventCreate(&es));
  for (auto &set : seg_evs)
    for (auto &ev : set) HIPCHK(h, hipEventCreate(&ev));
  HIPCHK(h, hipEventRecord(e0, h->stream));
  HIPCHK(h, hipMemsetAsync(badop, 0, 4, h->stream));
  const bool cand_kept = h->cand.valid && h->cand.ml == ml && h->cand.bytes_model == cfg->bytes_model &&
                         h->cand.lo_mask == lo_mask;
  if (!cand_kept) {
    h->cand.valid = h->cand.fused_built = false;
    hipLaunchKernelGGL(k_cand_words, dim3(blocks_for(nc, 256)), dim3(256), 0, h->stream, cwords, nc);
    hipLaunchKernelGGL(k_iota, dim3(blocks_for(nc, 256)), dim3(256), 0, h->stream, rank, nc);
  }
  // fingerprints of single-instruction slots, then of pairs
  if (!cand_kept) {
    uint32_t *w1;
    ENSURE(B_SCRATCH, uint32_t, kNumInsts, w1);
    hipLaunchKernelGGL(k_inst_words, dim3(blocks_for(kNumInsts, 256)), dim3(256), 0, h->stream, w1);
    hipLaunchKernelGGL(k_fingerprints, dim3(blocks_for(kNumInsts, 4 * kFpPerWave)), dim3(256), 0, h->stream, w1, 1u,
                       (uint64_t)kNumInsts, cfg->bytes_model, cfp, cmh, nullptr, lo_mask, cbytes, ccyc, cprune,
                       badop);
    if (ml == 2) {
      // pairs: words at cwords[2*slot] for slot >= 2863 are contiguous pairs
      hipLaunchKernelGGL(k_fingerprints, dim3(blocks_for(nc - kNumInsts, 4 * kFpPerWave)), dim3(256), 0, h->stream,
                         cwords + 2 * kNumInsts, 2u, nc - kNumInsts, cfg->bytes_model,
                         cfp + (uint64_t)kNumInsts * kFpVecs, cmh + kNumInsts, nullptr, lo_mask,
                         cbytes + kNumInsts, ccyc + kNumInsts, cprune + kNumInsts, badop);
    }
    HIPCHK(h, hipGetLastError());
  }
  // the QuickCheck hash join's candidate side: the filter hashes under the
  // masked compare, sorted, with a directory on their top dk bits (about one
  // candidate per bucket)
  {
    uint32_t *key, *val, *shs, *ssl, *dir;
    uint32_t dk = 1;
    while (dk < 24 && (1ull << dk) < nc) dk++;
    ENSURE(B_QJ_KEY, uint32_t, nc, key);
    ENSURE(B_QJ_VAL, uint32_t, nc, val);
    ENSURE(B_QJ_HASH, uint32_t, nc, shs);
    ENSURE(B_QJ_SLOT, uint32_t, nc, ssl);
    ENSURE(B_QJ_DIR, uint32_t, (1ull << dk) + 1, dir);
    h->qj_hash = shs;
    h->qj_slot = ssl;
    h->qj_dir = dir;
    h->qj_dk = dk;
  }
  if (!cand_kept) {
    uint32_t *key = (uint32_t *)h->bufs[B_QJ_KEY].p, *val = (uint32_t *)h->bufs[B_QJ_VAL].p;
    uint32_t *shs = (uint32_t *)h->qj_hash, *ssl = (uint32_t *)h->qj_slot, *dir = (uint32_t *)h->qj_dir;
    const uint32_t dk = h->qj_dk;
    hipLaunchKernelGGL(k_cand_filter_hash, dim3(blocks_for(nc, 256)), dim3(256), 0, h->stream, cfp, nc, lo_mask, key,
                       val);
    HIPCHK(h, hipGetLastError());
    size_t tmp_bytes = 0;
    HIPCHK(h, rocprim::radix_sort_pairs(nullptr, tmp_bytes, key, shs, val, ssl, (uint32_t)nc, 0, 32, h->stream));
    uint8_t *tmp;
    ENSURE(B_GSORT, uint8_t, tmp_bytes + 1, tmp);
    HIPCHK(h, rocprim::radix_sort_pairs(tmp, tmp_bytes, key, shs, val, ssl, (uint32_t)nc, 0, 32, h->stream));
    hipLaunchKernelGGL(k_hash_dir, dim3(blocks_for((1ull << dk) + 1, 256)), dim3(256), 0, h->stream, shs,
                       (uint32_t)nc, dk, dir);
    HIPCHK(h, hipGetLastError());
    h->cand.valid = true;
    h->cand.ml = ml;
    h->cand.bytes_model = cfg->bytes_model;
    h->cand.lo_mask = lo_mask;
  }
  // QuickCheck and MidCheck as one join (MidCheck on): the candidates' join
  // keys sorted with a directory, and their classes of equal masked
  // QuickCheck outputs (keys, first members, counts); a class whose members
  // differ (a 64-bit key collision) sends this enumeration to the separate
  // kernels
  // (a caller's QuickCheck tile asks for the all-pairs grid kernel)
  bool fused = cfg->use_mid && h->qcm_join && cfg->tile_targets == 0;
  unsigned long long *cls_key = nullptr;
  uint64_t *cls_n = nullptr;
  uint32_t *cls_rep = nullptr, *cls_cnt = nullptr, *j_key = nullptr, *j_slot = nullptr, *j_dir = nullptr, j_dk = 1;
  uint32_t *cls_dir = nullptr, cls_dk = 1;  // the classes' bucket directory on the top cls_dk key bits
  uint32_t *ccls_id = nullptr;              // each candidate's class
  uint2 *cps_sig = nullptr;                 // each candidate's P bytes on the QuickCheck vectors
  if (fused) {
    while (j_dk < 24 && (1ull << j_dk) < nc) j_dk++;
    unsigned long long *ck, *cks;
    uint32_t *v1, *v1s, *jk, *v2, *head, *bad;
    uint64_t *hoff;
    ENSURE(B_CK, unsigned long long, nc, ck);
    ENSURE(B_CKS, unsigned long long, nc, cks);
    ENSURE(B_CV1, uint32_t, nc, v1);
    ENSURE(B_CV1S, uint32_t, nc, v1s);
    ENSURE(B_JK, uint32_t, nc, jk);
    ENSURE(B_JV, uint32_t, nc, v2);
    ENSURE(B_JKS, uint32_t, nc, j_key);
    ENSURE(B_JVS, uint32_t, nc, j_slot);
    ENSURE(B_JDIR, uint32_t, (1ull << j_dk) + 1, j_dir);
    ENSURE(B_CHEAD, uint32_t, nc, head);
    ENSURE(B_CHOFF, uint64_t, nc + 1, hoff);
    ENSURE(B_CLSKEY, unsigned long long, nc, cls_key);
    ENSURE(B_CLSREP, uint32_t, nc, cls_rep);
    ENSURE(B_CLSCNT, uint32_t, nc * (kClsBins + 1), cls_cnt);
    ENSURE(B_CLSN, uint64_t, 2, cls_n);
    ENSURE(B_CCLS, uint32_t, nc, ccls_id);
    ENSURE(B_CPSIG, uint2, nc, cps_sig);
    while (cls_dk < 24 && (1ull << cls_dk) < nc) cls_dk++;
    ENSURE(B_CLSDIR, uint32_t, (1ull << cls_dk) + 1, cls_dir);
    bad = reinterpret_cast<uint32_t *>(cls_n + 1);
    const uint32_t n32 = (uint32_t)nc;
    if (h->cand.fused_built) {
      fused = h->cand.fused_ok;
    } else {
      hipLaunchKernelGGL(k_cand_keys, dim3(blocks_for(nc, 256)), dim3(256), 0, h->stream, cfp, cmh, nc, lo_mask, ck, jk,
                         v1, v2, cps_sig);
      HIPCHK(h, hipGetLastError());
      size_t tb1 = 0, tb2 = 0;
      HIPCHK(h, rocprim::radix_sort_pairs(nullptr, tb1, ck, cks, v1, v1s, n32, 0, 64, h->stream));
      HIPCHK(h, rocprim::radix_sort_pairs(nullptr, tb2, jk, j_key, v2, j_slot, n32, 0, 32, h->stream));
      uint8_t *tmp;
      ENSURE(B_GSORT, uint8_t, std::max(tb1, tb2) + 1, tmp);
      HIPCHK(h, rocprim::radix_sort_pairs(tmp, tb1, ck, cks, v1, v1s, n32, 0, 64, h->stream));
      HIPCHK(h, rocprim::radix_sort_pairs(tmp, tb2, jk, j_key, v2, j_slot, n32, 0, 32, h->stream));
      hipLaunchKernelGGL(k_hash_dir, dim3(blocks_for((1ull << j_dk) + 1, 256)), dim3(256), 0, h->stream, j_key, n32,
                         j_dk,
                         j_dir);
      hipLaunchKernelGGL(k_class_heads, dim3(blocks_for(nc, 256)), dim3(256), 0, h->stream, cks, n32, head);
      HIPCHK(h, hipGetLastError());
      int rc = scan_u32(h, head, nc, hoff, cls_n);
      if (rc) return rc;
      hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(nc * (kClsBins + 1), 256)), dim3(256), 0, h->stream, cls_cnt,
                         nc * (kClsBins + 1), 0u);
      hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, h->stream, bad, (uint64_t)1, 0u);
      hipLaunchKernelGGL(k_class_fill, dim3(blocks_for(nc, 256)), dim3(256), 0, h->stream, cks, v1s, n32, head, hoff,
                         cbytes, cprune, cls_key, cls_rep, cls_cnt, ccls_id);
      hipLaunchKernelGGL(k_class_verify, dim3(blocks_for(nc, 256)), dim3(256), 0, h->stream, v1s, n32, head, hoff,
                         cls_rep, cfp, lo_mask, bad);
      hipLaunchKernelGGL(k_class_dir, dim3(blocks_for((1ull << cls_dk) + 1, 256)), dim3(256), 0, h->stream, cls_key,
                         cls_n, cls_dk, cls_dir);
      HIPCHK(h, hipGetLastError());
      uint32_t bad_h = 0;
      HIPCHK(h, hipMemcpyAsync(&bad_h, bad, 4, hipMemcpyDeviceToHost, h->stream));
      SYNC_WD("device sync");
      if (bad_h) {
        trace(h, "  candidate classes: a 64-bit key holds two outputs; QuickCheck and MidCheck run separately");
        fused = false;
      }
      h->cand.fused_built = true;
      h->cand.fused_ok = fused;
    }
  }
  // the empty candidate's fingerprint: a length-0 sequence's outputs are the
  // vectors themselves (no word is read)
  uint2 *efp = nullptr;
  if (cfg->include_empty) {
    uint8_t *eb;
    ENSURE(B_EFP, uint8_t, 256, eb);
    efp = reinterpret_cast<uint2 *>(eb);  // [0, 8): fingerprint, [8]: Mid hashes, then bytes and cycles
    hipLaunchKernelGGL(k_fingerprints, dim3(1), dim3(256), 0, h->stream, cwords, 0u, (uint64_t)1, cfg->bytes_model,
                       efp, efp + 8, (uint32_t *)nullptr, lo_mask, eb + 80, eb + 81, (uint8_t *)nullptr, badop);
    HIPCHK(h, hipGetLastError());
  }
  // plans by opcode for length-2 targets (whose candidates are single instructions)
  // (a function of the ISA, the sweep mode and the mask only, so the handle
  // keeps it for the next call with the same two)
  PlanEntry *ptab = nullptr;
  if (tl == 2 && ml == 1) {
    ENSURE(B_PLANTAB, PlanEntry, kPlanTab, ptab);  // fixed size: allocated once
    if (h->ptab_mode != cfg->exh_mode || h->ptab_pm != pm_masked) {
      hipLaunchKernelGGL(k_plan_table, dim3(blocks_for(kPlanTab, 256)), dim3(256), 0, h->stream, cfg->exh_mode,
                         pm_masked, ptab);
      HIPCHK(h, hipGetLastError());
      h->ptab_mode = cfg->exh_mode;
      h->ptab_pm = pm_masked;
    }
  }

  // ---- segments of targets
  const uint64_t seg_max = 1ull << h->seg_log;
  const uint64_t seg = cfg->targets ? std::min<uint64_t>(nt_all, seg_max) : seg_max;
  // QuickCheck launch shape: the caller's tile (grid form), else the autotuner's choice
  // (a caller's tile is rounded up to whole 32-target scan steps: tiles start
  // on hash-block boundaries; the tile only distributes the work)
  sx_handle::QcChoice qcc{tl, ml, 0, (cfg->tile_targets + 4 * kQcBlk - 1) / (4 * kQcBlk) * (4 * kQcBlk), 0};
  bool qc_tuned = cfg->tile_targets != 0;
  // survivor buffer: the caller's size, else what sufficed before on this
  // handle, else about one entry per 4,096 pairs (QuickCheck passes ~1e-4 of
  // the pairs at --dead-flags 0 and a few times more when flags are dead)
  uint32_t surv_cap = cfg->survivor_cap;
  if (!surv_cap) {
    const uint64_t est = std::min<uint64_t>(seg, nt_all) * nc / 4096;
    surv_cap = (uint32_t)std::max<uint64_t>({1u << 25, std::min<uint64_t>(est, 1u << 28), h->surv_cap_hint});
  }
  // pm_masked (the masked compare's P mask) is defined with the candidates
  float ms_qc = 0, ms_mid = 0, ms_exh = 0, ms_setup = 0;
  HIPCHK(h, hipEventRecord(es, h->stream));  // end of the setup; read after the last sync
  auto harvest_stage_times = [&](int set) -> int {
    if (!seg_evs_pending[set]) return SX_OK;
    float a = 0, b = 0, c = 0;
    HIPCHK(h, hipEventSynchronize(seg_evs[set][3]));
    HIPCHK(h, hipEventElapsedTime(&a, seg_evs[set][0], seg_evs[set][1]));
    HIPCHK(h, hipEventElapsedTime(&b, seg_evs[set][1], seg_evs[set][2]));
    HIPCHK(h, hipEventElapsedTime(&c, seg_evs[set][2], seg_evs[set][3]));
    ms_qc += a;
    ms_mid += b;
    ms_exh += c;
    seg_evs_pending[set] = false;
    return SX_OK;
  };

  uint64_t seg_no = 0;
  PendingRules pend{0, 0, 0, 0};  // the last segment's rules, not yet queued for the host
  for (uint64_t s0 = 0; s0 < nt_all; s0 += seg, seg_no++) {
    uint64_t nt = std::min<uint64_t>(seg, nt_all - s0);
    const int evset = (int)(seg_no & 1);
    if (pend.n) {  // the previous segment's (its last kernels are still running)
      int rc = queue_pieces(h, pend);
      if (rc) return rc;
    }
    {
      int rc = harvest_stage_times(evset);
      if (rc) return rc;
    }
    hipEvent_t e1 = seg_evs[evset][0], e2 = seg_evs[evset][1], e3 = seg_evs[evset][2], e4 = seg_evs[evset][3];
    uint32_t *traw = nullptr;  // implicit length 3: offset of each target in the raw range
    if (implicit3) {
      uint32_t *tw3, *counts;
      uint64_t *offs, *total;
      const uint64_t nb = blocks_for(nt, 256);
      ENSURE(B_TARGETS, uint32_t, nt * 3, tw3);
      ENSURE(B_TRAW, uint32_t, nt, traw);
      ENSURE(B_GEN_COUNTS, uint32_t, nb, counts);
      ENSURE(B_GEN_OFFS, uint64_t, nb, offs);
      ENSURE(B_TOTAL, uint64_t, 1, total);
      hipLaunchKernelGGL(k_gen3_count, dim3((unsigned)nb), dim3(256), 0, h->stream, tbase + s0, (uint32_t)nt, counts);
      int rc = scan_u32(h, counts, nb, offs, total);
      if (rc) return rc;
      hipLaunchKernelGGL(k_gen3_write, dim3((unsigned)nb), dim3(256), 0, h->stream, tbase + s0, (uint32_t)nt, offs,
                         tw3, traw);
      HIPCHK(h, hipGetLastError());
      uint64_t kept = 0;
      HIPCHK(h, hipMemcpyAsync(&kept, total, 8, hipMemcpyDeviceToHost, h->stream));
      SYNC_WD("device sync");
      nt = kept;
      if (nt == 0) continue;
    }
    uint32_t *tw;
    uint2 *tfp, *tmh;
    uint32_t *tv0;  // the targets' QuickCheck filter hashes
    uint8_t *tbytes, *tcyc;
    ENSURE(B_TARGETS, uint32_t, nt * tl, tw);
    ENSURE(B_TFP, uint2, nt * kFpVecs, tfp);
    ENSURE(B_TMH, uint2, nt, tmh);
    ENSURE(B_TV0, uint32_t, nt + kQcBlk, tv0);  // slack for the block loads of QuickCheck
    ENSURE(B_TBYTES, uint8_t, nt, tbytes);
    ENSURE(B_TCYC, uint8_t, nt, tcyc);
    HIPCHK(h, hipEventRecord(e1, h->stream));
    if (cfg->targets) {
      HIPCHK(h, hipMemcpyAsync(tw, cfg->targets + s0 * tl, nt * tl * 4, hipMemcpyHostToDevice, h->stream));
    } else if (implicit3) {
      // already generated in place
    } else {
      HIPCHK(h, hipMemcpyAsync(tw, d_tsrc + s0 * 2, nt * 8, hipMemcpyDeviceToDevice, h->stream));
    }
    // the joined QuickCheck + MidCheck needs the Mid hashes only of targets
    // whose masked QuickCheck outputs equal some candidate's (about 15% of
    // the length-2 targets against single instructions at --dead-flags
    // 0x00): their QuickCheck outputs first, eight targets per pass, the Mid
    // hashes after the class lookup.  Against the 8.2M candidates of length
    // 1 and 2 nearly every target has a class, and the list only costs; the
    // empty candidate reads every target's Mid hashes.
    const bool fp_split = fused && h->fp_split && ml == 1 && !cfg->include_empty;
    // QuickCheck's counters (2 per shard) and, with fp_split, the class
    // members' list and its length (the last slot)
    unsigned long long *qsums = nullptr;
    uint32_t *tlist = nullptr, *tlist_n = nullptr, *tcls = nullptr;
    if (fused) {
      ENSURE(B_TCLS, uint32_t, nt, tcls);
      // [0, 2 * shards): counters; then the list length; then the gathered region counts
      ENSURE(B_QCSUM, unsigned long long, 2 * kQcCntShards + 1 + kMidRegions / 2, qsums);
      HIPCHK(h, hipMemsetAsync(qsums, 0, (2 * kQcCntShards + 1) * 8, h->stream));
    }
    if (fp_split) {
      ENSURE(B_TLIST, uint32_t, nt, tlist);
      tlist_n = reinterpret_cast<uint32_t *>(qsums + 2 * kQcCntShards);
      const QcClassArgs ca{cls_key, cls_dir, cls_dk, cls_rep, cls_cnt, cfp, lo_mask, qsums, tlist, tlist_n, tcls};
      size_t ts = 0;
      int rc = timed_begin(h, 0, &ts);
      if (rc) return rc;
      hipLaunchKernelGGL(k_fp_qc<true>, dim3(blocks_for(nt, 4 * fpq_per_wave((uint32_t)tl))), dim3(256), 0, h->stream,
                         tw, (uint32_t)tl, nt, cfg->bytes_model, tfp, tbytes, tcyc, nullptr, badop, ca);
      HIPCHK(h, hipGetLastError());
      rc = timed_end(h, ts);
      if (rc) return rc;
    } else {
      hipLaunchKernelGGL(k_fingerprints, dim3(blocks_for(nt, 4 * kFpPerWave)), dim3(256), 0, h->stream, tw,
                         (uint32_t)tl, nt, cfg->bytes_model, tfp, tmh, tv0, lo_mask, tbytes, tcyc, nullptr, badop);
    }
    HIPCHK(h, hipGetLastError());

    // the rounds' per-(target, pass) state, cleared while the host waits for
    // the MidCheck count
    uint32_t *cur, *rej, *rule_flag;  // rule_flag (first hit): slot k = target * 2 + pass - 1 holds a rule
    uint8_t *done;
    ENSURE(B_CUR, uint32_t, nt * 2, cur);
    ENSURE(B_REJ, uint32_t, nt * 2, rej);
    ENSURE(B_DONE, uint8_t, nt * 2, done);
    ENSURE(B_RULEFLAG, uint32_t, nt * 2, rule_flag);
    uint2 *mid;
    uint32_t nmid = 0;
    MidList mlist{};  // where the nmid survivors are in mid
    uint32_t *counters;  // [0, 8): MidCheck and the rounds; [8, 16): QuickCheck's region counts
    ENSURE(B_COUNTERS, uint32_t, 8 + kQcRegions, counters);
    if (fused) {
      // ---- QuickCheck's counters from the classes, then QuickCheck and
      // MidCheck as one join (re-run with a larger buffer on overflow)
      {
        size_t ts = 0;
        int rc = timed_begin(h, 0, &ts);
        if (rc) return rc;
        if (!fp_split)  // (with fp_split, k_fp_qc looked the classes up)
          hipLaunchKernelGGL(k_qc_count, dim3(blocks_for(nt, 256 * kQcCntChunks)), dim3(256), 0, h->stream, tfp,
                             tbytes, (uint32_t)nt, lo_mask, cls_key, cls_dir, cls_dk, cls_rep, cls_cnt, cfp, qsums,
                             (uint32_t *)nullptr, (uint32_t *)nullptr, tcls);
        HIPCHK(h, hipGetLastError());
        if (fp_split)
          hipLaunchKernelGGL(k_fp_mid_list, dim3(std::min(blocks_for(nt, 4 * kFpPerWave), 8u * h->num_cu)), dim3(256), 0,
                             h->stream, tw, tlist, tlist_n, (uint32_t)tl, tmh, lo_mask, badop);
        HIPCHK(h, hipGetLastError());
        rc = timed_end(h, ts);
        if (rc) return rc;
      }
      HIPCHK(h, hipEventRecord(e2, h->stream));
      // the caller's survivor buffer size, else what sufficed before, else
      // 2^22, over kMidRegions regions
      const uint32_t mid_total = cfg->survivor_cap ? cfg->survivor_cap : std::max<uint32_t>(h->mid_cap_hint, 1u << 22);
      uint32_t mid_cap = (mid_total + kMidRegions - 1) / kMidRegions;  // per region
      uint32_t *mcnt;
      ENSURE(B_MIDCNT, uint32_t, kMidRegions * kMidCntStride, mcnt);
      for (;;) {
        ENSURE(B_MID, uint2, (uint64_t)mid_cap * kMidRegions, mid);
        HIPCHK(h, hipMemsetAsync(counters, 0, 8 * 4, h->stream));  // counters[2..7] stay 0 for the rounds
        HIPCHK(h, hipMemsetAsync(mcnt, 0, kMidRegions * kMidCntStride * 4, h->stream));
        hipLaunchKernelGGL(k_seg_clear, dim3(blocks_for(nt * 2, 256)), dim3(256), 0, h->stream, cur, rej, done,
                           cfg->hits ? nullptr : rule_flag, nt * 2);
        size_t ts = 0;
        int rc = timed_begin(h, 0, &ts);
        if (rc) return rc;
        // waves per 64-target chunk: about 32 waves per CU in all
        const uint32_t chunks = blocks_for(nt, kWave);
        const uint32_t parts = std::min(16u, std::max(1u, 32u * (uint32_t)h->num_cu / std::max(chunks, 1u)));
        hipLaunchKernelGGL(k_qcm_join,
                           dim3(fp_split ? std::min(blocks_for(nt, 256), 16u * h->num_cu) : blocks_for(nt, 256), parts),
                           dim3(256), 0, h->stream, tfp, tmh, tbytes, (uint32_t)nt, lo_mask, j_key, j_slot, j_dir,
                           j_dk, cfp, cmh, cbytes, cprune, cfg->hits, mid, mid_cap, mcnt, rank,
                           cfg->hits ? nullptr : cur, tlist, tlist_n, tcls, ccls_id, cps_sig);
        HIPCHK(h, hipGetLastError());
        rc = timed_end(h, ts);
        if (rc) return rc;
        uint32_t *gathered = reinterpret_cast<uint32_t *>(qsums + 2 * kQcCntShards + 1);
        hipLaunchKernelGGL(k_gather_counts, dim3(1), dim3(64), 0, h->stream, mcnt, gathered);
        HIPCHK(h, hipGetLastError());
        struct {
          unsigned long long q[2 * kQcCntShards + 1 + kMidRegions / 2];
        } rb;
        rc = read_back(h, {{rb.q, qsums, sizeof(rb.q)}}, "QuickCheck and MidCheck");
        if (rc) return rc;
        const uint32_t *rbc = reinterpret_cast<const uint32_t *>(rb.q + 2 * kQcCntShards + 1);
        uint64_t total = 0;
        uint32_t most = 0;
        for (uint32_t r = 0; r < kMidRegions; r++) {
          const uint32_t nr = rbc[r];
          mlist.pre[r] = (uint32_t)std::min<uint64_t>(total, 0xFFFFFFFFull);
          total += nr;
          most = std::max(most, nr);
        }
        if (total > 0xFFFFFFFFull) return set_err(h, SX_ENOMEM, "MidCheck survivors (%llu)", (unsigned long long)total);
        mlist.pre[kMidRegions] = (uint32_t)total;
        mlist.cap = mid_cap;
        nmid = (uint32_t)total;
        if (most > mid_cap) {
          st.overflow_retries++;
          if ((uint64_t)most * 5 / 4 + 1024 > 0xFFFFFFFFull / kMidRegions)
            return set_err(h, SX_ENOMEM, "MidCheck survivor buffer (%u in one region)", most);
          mid_cap = most + most / 4 + 1024;
          if (!cfg->survivor_cap) h->mid_cap_hint = std::max(h->mid_cap_hint, mid_cap * kMidRegions);
          continue;
        }
        uint64_t raw = 0, hits = 0;
        for (uint32_t k = 0; k < kQcCntShards; k++) {
          raw += rb.q[2 * k];
          hits += rb.q[2 * k + 1];
        }
        st.qc_raw_hits += raw;
        st.qc_hits += hits;
        trace(h, "segment %llu: %llu targets, QuickCheck survivors %llu (%llu after filters), MidCheck %u (joined)",
              (unsigned long long)s0, (unsigned long long)nt, (unsigned long long)raw, (unsigned long long)hits,
              nmid);
        break;
      }
      st.qc_pairs += nt * (uint64_t)(ml == 1 ? kNumInsts : nc);
    } else {
      // ---- QuickCheck, re-run with a larger buffer on overflow (no hit is dropped)
      uint2 *surv;
      uint32_t *qc_counts = counters + 8;
      uint32_t nsurv = 0;
      SurvRegions regions;
      for (;;) {
        ENSURE(B_SURV, uint2, surv_cap, surv);
        if (!qc_tuned) {
          int rc = tune_qc(h, tl, ml, tv0, tfp, nt, cfp, nc, pm_masked, surv, surv_cap, qc_counts, &qcc);
          if (rc) return rc;
          qc_tuned = true;
        }
        HIPCHK(h, hipMemsetAsync(qc_counts, 0, kQcRegions * 4, h->stream));
        {
          size_t ts = 0;
          int rc = timed_begin(h, 0, &ts);
          if (rc) return rc;
          rc = launch_qc(h, qcc.tile, qcc.blocks_per_cu, tv0, tfp, nt, cfp, nc, pm_masked, surv, surv_cap,
                         qc_counts);
          if (rc) return rc;
          rc = timed_end(h, ts);
          if (rc) return rc;
        }
        uint32_t cnt8[kQcRegions];
        {
          int rc = read_back(h, {{cnt8, qc_counts, sizeof(cnt8)}}, "QuickCheck");
          if (rc) return rc;
        }
        const uint32_t region_cap = surv_cap / kQcRegions;
        uint64_t total = 0;
        uint32_t most = 0;
        regions.cap = region_cap;
        for (uint32_t r = 0; r < kQcRegions; r++) {
          regions.pre[r] = (uint32_t)std::min<uint64_t>(total, 0xFFFFFFFFu);
          total += cnt8[r];
          most = std::max(most, cnt8[r]);
        }
        regions.pre[kQcRegions] = (uint32_t)std::min<uint64_t>(total, 0xFFFFFFFFu);
        trace(h, "segment %llu: %llu targets, QuickCheck survivors %llu (largest region %u of %u)",
              (unsigned long long)s0, (unsigned long long)nt, (unsigned long long)total, most, region_cap);
        if (most <= region_cap) {
          nsurv = (uint32_t)total;
          break;
        }
        st.overflow_retries++;
        if ((uint64_t)most * 5 / 4 + 1024 > 0xFFFFFFFFull / kQcRegions)
          return set_err(h, SX_ENOMEM, "QuickCheck survivor buffer (%u in one region)", most);
        surv_cap = kQcRegions * (most + most / 4 + 1024);
        if (!cfg->survivor_cap) h->surv_cap_hint = std::max(h->surv_cap_hint, surv_cap);
      }
      HIPCHK(h, hipEventRecord(e2, h->stream));
      st.qc_raw_hits += nsurv;
      st.qc_pairs += nt * (uint64_t)(ml == 1 ? kNumInsts : nc);

      // ---- MidCheck
      for (;;) {
        ENSURE(B_MID, uint2, std::max<uint32_t>(nsurv, 1), mid);
        HIPCHK(h, hipMemsetAsync(counters, 0, 8 * 4, h->stream));  // counters[2..7] stay 0 for the rounds
        hipLaunchKernelGGL(k_seg_clear, dim3(blocks_for(nt * 2, 256)), dim3(256), 0, h->stream, cur, rej, done,
                           cfg->hits ? nullptr : rule_flag, nt * 2);
        // first hit: MidCheck also leaves the first round's minimum per (target, pass) in cur
        if (nsurv)
          hipLaunchKernelGGL(k_midcheck, dim3(blocks_for(nsurv, kBlk)), dim3(kBlk), 0, h->stream, surv, regions,
                             nsurv, tmh, tbytes, cmh, cbytes, cprune, cfg->hits, cfg->use_mid, mid,
                             std::max<uint32_t>(nsurv, 1), counters, rank, cfg->hits ? nullptr : cur);
        HIPCHK(h, hipGetLastError());
        uint32_t cnt[2];
        {
          int rc = read_back(h, {{cnt, counters, 8}}, "MidCheck");
          if (rc) return rc;
        }
        st.qc_hits += cnt[0];
        nmid = cnt[1];
        for (uint32_t r = 0; r <= kMidRegions; r++) mlist.pre[r] = r ? nmid : 0u;  // one list
        mlist.cap = std::max<uint32_t>(nmid, 1);
        trace(h, "  MidCheck: %u QC hits after filters, %u pass", cnt[0], cnt[1]);
        break;
      }
    }
    st.mid_pass += nmid;
    HIPCHK(h, hipEventRecord(e3, h->stream));

    // ---- ExhaustiveCheck (rounds for first-hit)
    Job *jobs;
    sx_rule *rules;
    // every round queues at most one job per open (target, pass) entry; the
    // first-hit tail queues the untested survivors of the open entries, and
    // is only chosen (below) while they number at most kBulkRatio x a round's
    // jobs or kBulkJobs; all hits, or a tail forced after a fixed round
    // (SX_BULK_AFTER), may queue every survivor
    constexpr uint32_t kBulkRatio = 4;
    constexpr uint32_t kBulkJobs = 1u << 22;
    // The tail's round: by default at once while the survivors number at most
    // the (target, pass) entries (C2 at 0x00: 5.4M survivors for 15.3M
    // entries; the batch sweeps 10% more pairs than a round and a tail, and
    // saves the round's launches and read-backs), else after one round (at
    // 0xFF the first round resolves most entries out of ~470M survivors).
    const int bulk_at = h->bulk_after == -2 ? (nmid <= 2 * nt ? 0 : 1) : h->bulk_after;
    uint64_t job_bound = nmid;
    if (!cfg->hits && bulk_at < 0)
      job_bound = std::min<uint64_t>(nmid, std::max<uint64_t>((uint64_t)kBulkRatio * 2 * nt, kBulkJobs));
    const uint32_t job_cap = (uint32_t)std::max<uint64_t>(job_bound, 1);
    ENSURE(B_JOBS, Job, job_cap, jobs);
    uint32_t *jchunks, *jfail;  // written by the job kernels, read by the sweep
    ENSURE(B_CHUNKS, uint32_t, job_cap + 1, jchunks);
    ENSURE(B_FAIL, uint32_t, job_cap + 1, jfail);
    const uint32_t rule_cap = cfg->hits ? std::max<uint32_t>(nmid, 1) : (uint32_t)(2 * nt);
    ENSURE(B_RULES, sx_rule, rule_cap, rules);
    uint32_t *nrules_d = counters + 2, *njobs_d = counters + 3;  // zeroed before MidCheck
    // The first-hit tail runs as one batch once the untested survivors of the
    // open (target, pass) entries number at most kBulkJobs, or at most
    // kBulkRatio x the entries themselves (one batch then costs at most that
    // many rounds' sweeps); or after a fixed number of rounds when
    // SX_BULK_AFTER is set.  Rounds stay in use while a batch would sweep far
    // more pairs than the first hits need (at --dead-flags 0xFF the first
    // round resolves most targets out of ~470M survivors).
    uint32_t *nrest_d = counters + 4;
    bool go_bulk = false;
    for (int round = 0; nmid > 0; round++) {
      // every round rejects or resolves one candidate per open (target, pass)
      if ((uint64_t)round > nc + 1) return set_err(h, SX_EHIP, "first-hit rounds did not converge");
      if (!cfg->hits && bulk_at >= 0 && round >= bulk_at) go_bulk = true;
      if (!cfg->hits && go_bulk) {
        // the batch's minimum: in cur (kRankNone after a round's resolve), or at
        // round 0, where cur still holds MidCheck's minima, as complements in
        // rej (all 0 until a round rejects)
        uint32_t *best = round == 0 ? rej : cur;
        HIPCHK(h, hipMemsetAsync(njobs_d, 0, 4, h->stream));
        uint32_t *bkey = nullptr, *bidx = nullptr;  // the jobs' group keys, written as they are planned
        if (h->exh_group) {
          ENSURE(B_GKEY, uint32_t, job_cap, bkey);
          ENSURE(B_GIDX, uint32_t, job_cap, bidx);
        }
        uint4 *jent;
        ENSURE(B_JENT, uint4, job_cap, jent);
        hipLaunchKernelGGL(k_bulk_jobs, dim3(blocks_for(nmid, kBlk)), dim3(kBlk), 0, h->stream, mid, mlist, rank,
                           rej, done, tw, (uint32_t)tl, pm_masked, jobs, jent, job_cap, njobs_d, cfg->exh_mode, ptab,
                           jchunks, jfail, bkey, bidx);
        HIPCHK(h, hipGetLastError());
        uint32_t nj = nmid;  // at round 0 nothing is resolved or rejected: every survivor is a job
        if (round > 0) {
          int rc = read_back(h, {{&nj, njobs_d, 4}}, "first-hit tail jobs");
          if (rc) return rc;
        }
        if (nj == 0) break;
        if (nj > job_cap) return set_err(h, SX_ENOMEM, "job buffer overflow (%u > %u)", nj, job_cap);
        trace(h, "  first-hit tail: %u jobs", nj);
        st.exh_rounds++;
        st.exh_pairs += nj;
        uint64_t chunks = 0;
        // with the plan table the keyed jobs are compact records only
        int rc = run_exhaustive(h, nj, cfg->exh_mode, &chunks, true, ptab, pm_masked, bkey != nullptr,
                                ptab ? jent : nullptr);
        if (rc) return rc;
        // cur holds kRankNone for every entry after a resolve: reuse it for the minimum
        hipLaunchKernelGGL(k_bulk_min, dim3(blocks_for(nj, 256)), dim3(256), 0, h->stream, (const uint4 *)jent, nj,
                           (const uint32_t *)h->bufs[B_FAIL].p, rank, tfp, cfp, best, round == 0 ? 1 : 0);
        hipLaunchKernelGGL(k_bulk_emit, dim3(blocks_for(nt * 2, 256)), dim3(256), 0, h->stream, best,
                           (uint32_t)nt, tw, (uint32_t)tl, tfp, cfp, tbytes, tcyc, cbytes, ccyc, tbase + s0, rules,
                           rule_flag, round == 0 ? 1 : 0);
        HIPCHK(h, hipGetLastError());
        break;
      }
      HIPCHK(h, hipMemsetAsync(njobs_d, 0, 8, h->stream));  // njobs and nrest
      const bool first_round = !cfg->hits && round == 0;
      uint32_t *rkey = nullptr, *ridx = nullptr;
      if (first_round) {
        // cur holds the minima (MidCheck); every survivor is untested
        hipLaunchKernelGGL(k_entry_jobs, dim3(blocks_for(nt * 2, kBlk)), dim3(kBlk), 0, h->stream, cur,
                           (uint32_t)(nt * 2), tw, (uint32_t)tl, pm_masked, jobs, job_cap, njobs_d);
      } else {
        if (!cfg->hits)
          hipLaunchKernelGGL(k_round_min, dim3(blocks_for(nmid, 256)), dim3(256), 0, h->stream, mid, mlist, rank,
                             cur, rej, done);
        // all hits: the jobs are planned (and keyed) here
        if (cfg->hits && h->exh_group) {
          ENSURE(B_GKEY, uint32_t, job_cap, rkey);
          ENSURE(B_GIDX, uint32_t, job_cap, ridx);
        }
        hipLaunchKernelGGL(k_round_jobs, dim3(blocks_for(nmid, kBlk)), dim3(kBlk), 0, h->stream, mid, mlist, rank,
                           cur, rej, done, tw, (uint32_t)tl, pm_masked, cfg->hits, jobs, job_cap, njobs_d, nrest_d,
                           cfg->hits ? 1 : 0, cfg->exh_mode, ptab, jchunks, jfail, rkey, ridx);
      }
      HIPCHK(h, hipGetLastError());
      uint32_t cnt2[2] = {0, 0};
      {
        int rc = read_back(h, {{cnt2, njobs_d, 8}}, "round jobs");
        if (rc) return rc;
      }
      if (first_round) cnt2[1] = nmid;
      uint32_t nj = cnt2[0];
      if (nj == 0) break;
      if (!cfg->hits && bulk_at < 0 &&
          ((uint64_t)cnt2[1] <= (uint64_t)kBulkRatio * nj || cnt2[1] <= kBulkJobs)) {
        // drop this round's jobs: the batch covers them; cur back to "none" for the
        // minimum (for the first round the batch branch does it)
        if (round > 0)
          hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(nt * 2, 256)), dim3(256), 0, h->stream, cur, nt * 2,
                             kRankNone);
        HIPCHK(h, hipGetLastError());
        go_bulk = true;
        round--;
        continue;
      }
      if (nj > job_cap) return set_err(h, SX_ENOMEM, "job buffer overflow (%u > %u)", nj, job_cap);
      trace(h, "  round %d: %u jobs", round, nj);
      st.exh_rounds++;
      st.exh_pairs += nj;
      uint64_t chunks = 0;
      int rc = run_exhaustive(h, nj, cfg->exh_mode, &chunks, cfg->hits != 0, ptab, pm_masked,
                              rkey != nullptr);  // first hit: planned (and keyed) there
      if (rc) return rc;
      hipLaunchKernelGGL(k_round_resolve, dim3(blocks_for(nj, 256)), dim3(256), 0, h->stream,
                         (const Job *)h->bufs[B_JOBS].p, nj, (const uint32_t *)h->bufs[B_FAIL].p, rank, cur, rej,
                         done, tfp, cfp, tbytes, tcyc, cbytes, ccyc, cfg->hits, cfg->dead_flags, tbase + s0,
                         rules, rule_cap, nrules_d, rule_flag);
      HIPCHK(h, hipGetLastError());
      if (cfg->hits) break;
    }
    if (ml_req == 3) {
      int rc = run_l3(h, (uint32_t)nt, tw, (uint32_t)tl, tfp, tbytes, tcyc, rule_flag, rules, cfg->dead_flags,
                      cfg->bytes_model, cfg->exh_mode, tbase + s0, &st);
      if (rc) return rc;
    }
    uint32_t nr = 0;
    const sx_rule *rules_out = rules;
    if (cfg->hits) {
      HIPCHK(h, hipMemcpyAsync(&nr, nrules_d, 4, hipMemcpyDeviceToHost, h->stream));
    } else {
      // the filled slots in (target, pass) order
      uint64_t *rsums, *rtot;
      sx_rule *packed;
      const uint64_t nk = nt * 2, nb = (nk + kScanTile - 1) / kScanTile;
      ENSURE(B_SUMS, uint64_t, nb + 1, rsums);
      ENSURE(B_RULETOT, uint64_t, 1, rtot);
      hipLaunchKernelGGL(k_tile_sums, dim3((unsigned)nb), dim3(256), 0, h->stream, rule_flag, nk, rsums, 1);
      hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, h->stream, rsums, nb, rtot);
      HIPCHK(h, hipGetLastError());
      uint64_t tot = 0;
      int rc = read_back(h, {{&tot, rtot, 8}}, "rule count");
      if (rc) return rc;
      nr = (uint32_t)tot;
      ENSURE(B_RULES2, sx_rule, std::max<uint32_t>(nr, 1), packed);
      hipLaunchKernelGGL(k_compact_rules, dim3((unsigned)nb), dim3(256), 0, h->stream, rules, rule_flag, nk,
                         (const uint64_t *)rsums, packed);
      HIPCHK(h, hipGetLastError());
      rules_out = packed;
    }
    HIPCHK(h, hipEventRecord(e4, h->stream));
    seg_evs_pending[evset] = true;
    if (cfg->hits) SYNC_WD("device sync");  // nr
    if (nr > rule_cap) return set_err(h, SX_ENOMEM, "rule buffer overflow (%u rules)", nr);
    if (implicit3 && nr) {
      // rules name their target by its raw Go-order triple index
      hipLaunchKernelGGL(k_fix_raw_index, dim3(blocks_for(nr, 256)), dim3(256), 0, h->stream,
                         const_cast<sx_rule *>(rules_out), nr, traw, tbase + s0);
      HIPCHK(h, hipGetLastError());
    }
    if (h->rules_on_device) {
      // append to the device-resident list (in 32-bit words), doubling its
      // capacity as needed
      const uint64_t rw = h->rec_words, need = (h->d_rules_n + nr) * rw, used = h->d_rules_n * rw;
      if (need > h->d_rules_capw) {
        uint64_t cap2 = std::max<uint64_t>(std::max<uint64_t>(h->d_rules_capw * 2, need), 12u << 20);
        uint32_t *grown = nullptr;
        if (hipMalloc(&grown, cap2 * 4) != hipSuccess) return set_err(h, SX_ENOMEM, "rule list");
        if (used)
          HIPCHK(h, hipMemcpyAsync(grown, h->d_rules, used * 4, hipMemcpyDeviceToDevice, h->stream));
        SYNC_WD("device sync");
        HIPCHK(h, hipStreamSynchronize(h->copy_stream));  // copies in flight read the old list
        HIPCHK(h, hipStreamSynchronize(h->copy_stream2));
        if (h->d_rules) HIPCHK(h, hipFree(h->d_rules));
        h->d_rules = grown;
        h->d_rules_capw = cap2;
      }
      if (need > h->h_rules_capw) {
        // grow the pinned host copy (kept across calls, so this happens once
        // per size class)
        HIPCHK(h, hipStreamSynchronize(h->copy_stream));
        HIPCHK(h, hipStreamSynchronize(h->copy_stream2));
        uint64_t cap2 = std::max<uint64_t>(std::max<uint64_t>(h->h_rules_capw * 2, need), 12u << 20);
        uint32_t *grown = nullptr;
        if (hipHostMalloc(&grown, cap2 * 4, hipHostMallocDefault) != hipSuccess) {
          (void)hipGetLastError();
          return set_err(h, SX_ENOMEM, "pinned rule list");
        }
        if (used) memcpy(grown, h->h_rules, used * 4);
        if (h->h_rules) HIPCHK(h, hipHostFree(h->h_rules));
        h->h_rules = grown;
        h->h_rules_capw = cap2;
      }
      if (nr) {
        if (rw == kPackedWords) {
          hipLaunchKernelGGL(k_pack_rules, dim3(blocks_for(nr, 256)), dim3(256), 0, h->stream, rules_out, nr,
                             h->d_rules + used);
          HIPCHK(h, hipGetLastError());
        } else {
          HIPCHK(h, hipMemcpyAsync(h->d_rules + used, rules_out, (size_t)nr * sizeof(sx_rule),
                                   hipMemcpyDeviceToDevice, h->stream));
        }
        // copy the segment's rules to the host while the next segments run:
        // the copy stream waits for the packing here; its copies are queued
        // after the next launches on the library's stream (the next
        // segment's first ones, or the final read-back's), so the device does
        // not idle while the host queues them
        HIPCHK(h, hipEventRecord(h->seg_ev, h->stream));
        HIPCHK(h, hipStreamWaitEvent(h->copy_stream, h->seg_ev, 0));
        HIPCHK(h, hipStreamWaitEvent(h->copy_stream2, h->seg_ev, 0));
        pend = {used, nr, h->d_rules_n, rw};
      }
      h->d_rules_n += nr;
    } else {
      size_t old = h->rules.size();
      h->rules.resize(old + nr);
      if (nr) {
        // on the library's stream: k_fix_raw_index above may still be
        // rewriting these rules (a plain hipMemcpy on the null stream would
        // not wait for this non-blocking stream)
        HIPCHK(h, hipMemcpyAsync(h->rules.data() + old, rules_out, nr * sizeof(sx_rule), hipMemcpyDeviceToHost,
                                 h->stream));
        SYNC_WD("rule copy");
      }
    }
    if (cfg->include_empty) {
      // the empty candidate: one slot per target, in a stream of its own
      st.qc_pairs += nt;
      Job *ej;
      uint32_t *ech, *efl, *ecnt = counters + 5;  // [5] jobs, [6] rules
      sx_rule *er;
      ENSURE(B_JOBS, Job, nt, ej);
      ENSURE(B_CHUNKS, uint32_t, nt + 1, ech);
      ENSURE(B_FAIL, uint32_t, nt + 1, efl);
      hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, h->stream, ecnt, (uint64_t)2, 0u);
      hipLaunchKernelGGL(k_empty_jobs, dim3(blocks_for(nt, 256)), dim3(256), 0, h->stream, tfp, tmh, efp, efp + 8,
                         (uint32_t)nt, tw, (uint32_t)tl, pm_masked, cfg->hits, cfg->dead_flags ? 1 : 0, cfg->use_mid,
                         cfg->exh_mode, ej, ecnt, ech, efl);
      HIPCHK(h, hipGetLastError());
      uint32_t nje = 0;
      HIPCHK(h, hipMemcpyAsync(&nje, ecnt, 4, hipMemcpyDeviceToHost, h->stream));
      SYNC_WD("device sync");
      if (nje) {
        st.exh_pairs += nje;
        uint64_t chunks = 0;
        int rc = run_exhaustive(h, nje, cfg->exh_mode, &chunks, true);
        if (rc) return rc;
        ENSURE(B_ERULES, sx_rule, nje, er);
        hipLaunchKernelGGL(k_empty_emit, dim3(blocks_for(nje, 256)), dim3(256), 0, h->stream,
                           (const Job *)h->bufs[B_JOBS].p, nje, (const uint32_t *)h->bufs[B_FAIL].p, tfp, efp,
                           tbytes, tcyc, cfg->hits, cfg->dead_flags, tbase + s0, traw, er, ecnt + 1);
        HIPCHK(h, hipGetLastError());
        uint32_t ner = 0;
        HIPCHK(h, hipMemcpyAsync(&ner, ecnt + 1, 4, hipMemcpyDeviceToHost, h->stream));
        SYNC_WD("device sync");
        const size_t old = h->empty_rules.size();
        h->empty_rules.resize(old + ner);
        if (ner) {
          HIPCHK(h, hipMemcpyAsync(h->empty_rules.data() + old, er, (size_t)ner * sizeof(sx_rule),
                                   hipMemcpyDeviceToHost, h->stream));
          SYNC_WD("empty-candidate rules");
        }
      }
    }
    st.targets += nt;
  }
  uint32_t bad = 0;
  {
    // synchronises; the last segment's rule copies are queued before the wait
    int rc = read_evaluated(h, &st.exh_states, &st.exh_group_states, &st.exh_group_jobs, &bad, badop, &pend);
    if (rc) return rc;
  }
  HIPCHK(h, hipEventSynchronize(es));  // long complete unless there were no targets
  HIPCHK(h, hipEventElapsedTime(&ms_setup, e0, es));
  for (int set = 0; set < 2; set++) {
    int rc = harvest_stage_times(set);
    if (rc) return rc;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(es);
  for (auto &set : seg_evs)
    for (auto &ev : set) (void)hipEventDestroy(ev);
  if (bad) return set_err(h, SX_EOPCODE, "opcode out of range in a target");
  // first hit: segments in order, slots compacted in (target, pass) order, so
  // the rules are already sorted; all hits: sort on the host
  if (cfg->hits) std::sort(h->rules.begin(), h->rules.end(), [](const sx_rule &x, const sx_rule &y) {
    if (x.target_index != y.target_index) return x.target_index < y.target_index;
    if (x.pass != y.pass) return x.pass < y.pass;
    if (x.rep_len != y.rep_len) return x.rep_len < y.rep_len;
    for (int q = 0; q < 4; q++)
      if (x.rep[q] != y.rep[q]) return x.rep[q] < y.rep[q];
    return false;
  });
  std::sort(h->empty_rules.begin(), h->empty_rules.end(), [](const sx_rule &x, const sx_rule &y) {
    return x.target_index != y.target_index ? x.target_index < y.target_index : x.pass < y.pass;
  });
  st.empty_rules = h->empty_rules.size();
  SYNC_WD("device sync");
  const uint64_t n_rules = h->rules_on_device ? h->d_rules_n : h->rules.size();
  st.rules = n_rules;
  st.ms_setup = ms_setup;
  st.qc_tile = qcc.tile;
  st.qc_blocks_per_cu = qcc.blocks_per_cu;
  st.ms_qc_kernel = h->k_qc_ms - qc_ms0;
  st.ms_exh_kernel = h->k_exh_ms - exh_ms0;
  st.qc_launches = h->k_qc_launches - qc_l0;
  st.exh_launches = h->k_exh_launches - exh_l0;
  st.ms_exh_group_kernel = h->k_grp_ms - grp_ms0;
  st.exh_group_launches = h->k_grp_launches - grp_l0;
  st.ms_qc = ms_qc;
  st.ms_mid = ms_mid;
  st.ms_exh = ms_exh;
  st.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
  if (stats) *stats = st;
  if (n_out) *n_out = n_rules;
  if (out) {
    if (cap < n_rules) return SX_ECAPACITY;
    uint64_t got = 0;
    return sx_fetch_rules(h, out, 0, n_rules, &got);
  }
  return SX_OK;
}

namespace {
// Rules [offset, offset + m) from the pinned host copy into out, piece by
// piece as the copy's events complete, on up to 16 threads (thread k takes
// pieces k, k + T, ...); a piece's part is unpacked (packed records) or
// copied.
int fetch_pieces(sx_handle *h, sx_rule *out, uint64_t offset, uint64_t m) {
  const uint64_t end = offset + m;
  // work items: slices of kFetchSlice rules of the pieces that overlap the
  // range, in piece order, taken by the threads one at a time, so the pieces
  // that land last are unpacked by all threads at once
  constexpr uint64_t kFetchSlice = 8192;
  struct Item {
    size_t piece;
    uint64_t a, b;
  };
  std::vector<Item> items;
  for (size_t i = 0; i < h->pieces.size(); i++) {
    const auto &pc = h->pieces[i];
    const uint64_t a = std::max(pc.first, offset), b = std::min(pc.first + pc.second, end);
    for (uint64_t x = a; x < b; x += kFetchSlice) items.push_back({i, x, std::min(b, x + kFetchSlice)});
  }
  const uint32_t rw = h->rec_words;
  std::atomic<int> err{0};
  std::atomic<size_t> next{0};
  auto work = [&]() {
    for (;;) {
      const size_t j = next.fetch_add(1);
      if (j >= items.size() || err) return;
      const Item &it = items[j];
      // poll: a blocking wait sleeps until an interrupt, far longer than a piece
      hipError_t e;
      while ((e = hipEventQuery(h->piece_ev[it.piece])) == hipErrorNotReady) std::this_thread::yield();
      if (e != hipSuccess) {
        err = 1;
        return;
      }
      if (rw == kPackedWords)
        unpack_rules(out + (it.a - offset), h->h_rules + it.a * kPackedWords, it.b - it.a, h->fetch_model);
      else
        memcpy(out + (it.a - offset), h->h_rules + it.a * 12, (it.b - it.a) * sizeof(sx_rule));
    }
  };
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const size_t nth = std::min<size_t>((unsigned)h->fetch_threads, hw);
  auto &P = h->fetch_pool;
  if (items.size() > 1 && P.th.size() + 1 < nth) {
    try {
      while (P.th.size() + 1 < nth)
        P.th.emplace_back([&P] {
          uint64_t seen = 0;
          std::unique_lock<std::mutex> lk(P.mu);
          for (;;) {
            P.cv.wait(lk, [&] { return P.stop || P.gen != seen; });
            if (P.stop) return;
            seen = P.gen;
            const std::function<void()> *job = P.job;
            lk.unlock();
            (*job)();
            lk.lock();
            if (--P.busy == 0) P.cv_done.notify_all();
          }
        });
    } catch (...) {  // no more threads to be had: the ones started and this one take every item
    }
  }
  const std::function<void()> job = work;
  const bool helpers = items.size() > 1 && !P.th.empty();
  if (helpers) {
    std::lock_guard<std::mutex> lk(P.mu);
    P.job = &job;
    P.busy = P.th.size();
    P.gen++;
    P.cv.notify_all();
  }
  work();
  if (helpers) {
    std::unique_lock<std::mutex> lk(P.mu);
    P.cv_done.wait(lk, [&] { return P.busy == 0; });
    P.job = nullptr;
  }
  if (err) return set_err(h, SX_EHIP, "rule copy to the host failed");
  return SX_OK;
}
}  // namespace

extern "C" int sx_fetch_rules(sx_handle *h, sx_rule *out, uint64_t offset, uint64_t cap, uint64_t *n_copied) {
  if (!h || (!out && cap)) return SX_EINVAL;
  const uint64_t n = h->rules_on_device ? h->d_rules_n : h->rules.size();
  uint64_t m = offset >= n ? 0 : std::min<uint64_t>(cap, n - offset);
  if (m) {
    if (h->rules_on_device) {
      HIPCHK(h, hipSetDevice(h->device));
      int rc = fetch_pieces(h, out, offset, m);
      if (rc) return rc;
    } else {
      memcpy(out, h->rules.data() + offset, m * sizeof(sx_rule));
    }
  }
  if (n_copied) *n_copied = m;
  return SX_OK;
}

extern "C" int sx_fetch_empty_rules(sx_handle *h, sx_rule *out, uint64_t offset, uint64_t cap,
                                    uint64_t *n_copied) {
  if (!h || (!out && cap)) return SX_EINVAL;
  const uint64_t n = h->empty_rules.size();
  const uint64_t m = offset >= n ? 0 : std::min<uint64_t>(cap, n - offset);
  if (m) memcpy(out, h->empty_rules.data() + offset, m * sizeof(sx_rule));
  if (n_copied) *n_copied = m;
  return SX_OK;
}

extern "C" int sx_check_pairs(sx_handle *h, const uint32_t *t_words, const uint8_t *t_len,
                              const uint32_t *c_words, const uint8_t *c_len, uint64_t n,
                              uint32_t stage_mask, uint32_t dead_flags, int32_t exh_mode, uint8_t *out) {
  if (!h || (n && (!t_words || !t_len || !c_words || !c_len || !out))) return SX_EINVAL;
  if (h->poisoned) return set_err(h, SX_EHIP, "handle unusable after a device timeout");
  if (dead_flags > 0xFF || exh_mode < 0 || exh_mode > 1 || (stage_mask & ~7u)) return set_err(h, SX_EINVAL, "bad argument");
  for (uint64_t i = 0; i < n; i++) {
    if (t_len[i] > 4 || c_len[i] > 4) return set_err(h, SX_EINVAL, "sequence longer than 4");
    for (int k = 0; k < t_len[i]; k++)
      if ((t_words[i * 4 + k] & 0xFFFFu) >= OP_COUNT) return set_err(h, SX_EOPCODE, "opcode out of range");
    for (int k = 0; k < c_len[i]; k++)
      if ((c_words[i * 4 + k] & 0xFFFFu) >= OP_COUNT) return set_err(h, SX_EOPCODE, "opcode out of range");
  }
  memset(out, 0, n);
  if (n == 0) return SX_OK;
  HIPCHK(h, hipSetDevice(h->device));
  // zero the words beyond each length so unused slots are inert
  std::vector<uint32_t> tw(t_words, t_words + 4 * n), cw(c_words, c_words + 4 * n);
  for (uint64_t i = 0; i < n; i++)
    for (int k = 0; k < 4; k++) {
      if (k >= t_len[i]) tw[i * 4 + k] = 0;
      if (k >= c_len[i]) cw[i * 4 + k] = 0;
    }
  uint32_t *dtw, *dcw;
  uint8_t *dtl, *dcl, *o1, *o2;
  ENSURE(B_P_TW, uint32_t, 4 * n, dtw);
  ENSURE(B_P_CW, uint32_t, 4 * n, dcw);
  ENSURE(B_P_TL, uint8_t, n, dtl);
  ENSURE(B_P_CL, uint8_t, n, dcl);
  ENSURE(B_P_OUT1, uint8_t, n, o1);
  ENSURE(B_P_OUT2, uint8_t, n, o2);
  HIPCHK(h, hipMemcpyAsync(dtw, tw.data(), 16 * n, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(dcw, cw.data(), 16 * n, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(dtl, t_len, n, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(dcl, c_len, n, hipMemcpyHostToDevice, h->stream));
  const uint32_t pm = kPMask & ~dead_flags;
  std::vector<uint8_t> q(n), m(n), e(n, 0);
  if (stage_mask & (SX_STAGE_QC | SX_STAGE_MID)) {
    hipLaunchKernelGGL(k_pair_vectors, dim3(blocks_for(n, 4)), dim3(256), 0, h->stream, dtw, dtl, dcw, dcl, n, pm,
                       o1, o2, (uint8_t *)nullptr);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipMemcpyAsync(q.data(), o1, n, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(m.data(), o2, n, hipMemcpyDeviceToHost, h->stream));
  }
  if (stage_mask & SX_STAGE_EXH) {
    const uint64_t batch = 1u << 22;
    for (uint64_t b0 = 0; b0 < n; b0 += batch) {
      uint32_t nb = (uint32_t)std::min<uint64_t>(batch, n - b0);
      std::vector<Job> jobs(nb);
      for (uint32_t i = 0; i < nb; i++) {
        Job &j = jobs[i];
        memset(&j, 0, sizeof(j));
        for (int k = 0; k < kJobWords; k++) {
          j.tw[k] = k < 4 ? tw[(b0 + i) * 4 + k] : 0u;
          j.cw[k] = k < 4 ? cw[(b0 + i) * 4 + k] : 0u;
        }
        j.tl = t_len[b0 + i];
        j.cl = c_len[b0 + i];
        j.pass = 1;
        j.pmask = (uint8_t)pm;
        j.target = i;
      }
      Job *dj;
      ENSURE(B_JOBS, Job, nb, dj);
      HIPCHK(h, hipMemcpyAsync(dj, jobs.data(), nb * sizeof(Job), hipMemcpyHostToDevice, h->stream));
      uint64_t chunks = 0;
      int rc = run_exhaustive(h, nb, exh_mode, &chunks);
      if (rc) return rc;
      std::vector<uint32_t> f(nb);
      HIPCHK(h, hipMemcpyAsync(f.data(), h->bufs[B_FAIL].p, nb * 4, hipMemcpyDeviceToHost, h->stream));
      SYNC_WD("device sync");
      for (uint32_t i = 0; i < nb; i++) e[b0 + i] = f[i] ? 0 : 1;
    }
  }
  SYNC_WD("device sync");
  for (uint64_t i = 0; i < n; i++) {
    uint8_t v = 0;
    if ((stage_mask & SX_STAGE_QC) && q[i]) v |= SX_STAGE_QC;
    if ((stage_mask & SX_STAGE_MID) && m[i]) v |= SX_STAGE_MID;
    if ((stage_mask & SX_STAGE_EXH) && e[i]) v |= SX_STAGE_EXH;
    out[i] = v;
  }
  return SX_OK;
}

// sx_check_pairs with every device buffer in the scratch array: no hipMalloc,
// hipFree or other allocation (host or device), for callers that cannot
// allocate.  Same arguments and result bits; n <= kNaMaxPairs.
extern "C" int sx_check_pairs_noalloc(sx_handle *h, const uint32_t *t_words, const uint8_t *t_len,
                                      const uint32_t *c_words, const uint8_t *c_len, uint64_t n,
                                      uint32_t stage_mask, uint32_t dead_flags, int32_t exh_mode, uint8_t *out) {
  if (!h || (n && (!t_words || !t_len || !c_words || !c_len || !out))) return SX_EINVAL;
  if (h->poisoned) return set_err(h, SX_EHIP, "handle unusable after a device timeout");
  if (dead_flags > 0xFF || exh_mode < 0 || exh_mode > 1 || (stage_mask & ~7u)) return set_err(h, SX_EINVAL, "bad argument");
  if (n > kNaMaxPairs)
    return set_err(h, SX_EINVAL, "sx_check_pairs_noalloc: %llu pairs exceed its maximum of %llu per call",
                   (unsigned long long)n, (unsigned long long)kNaMaxPairs);
  if (!h->scratch)
    return set_err(h, SX_EHIP, "scratch array address unavailable: %s", hipGetErrorString(h->scratch_rc));
  for (uint64_t i = 0; i < n; i++) {
    if (t_len[i] > 4 || c_len[i] > 4) return set_err(h, SX_EINVAL, "sequence longer than 4");
    for (int k = 0; k < t_len[i]; k++)
      if ((t_words[i * 4 + k] & 0xFFFFu) >= OP_COUNT) return set_err(h, SX_EOPCODE, "opcode out of range");
    for (int k = 0; k < c_len[i]; k++)
      if ((c_words[i * 4 + k] & 0xFFFFu) >= OP_COUNT) return set_err(h, SX_EOPCODE, "opcode out of range");
  }
  if (n == 0) return SX_OK;
  HIPCHK(h, hipSetDevice(h->device));
  uint8_t *const S = h->scratch;  // never the array's own name (see above)
  uint32_t *dtw = reinterpret_cast<uint32_t *>(S + kNa.tw), *dcw = reinterpret_cast<uint32_t *>(S + kNa.cw);
  uint8_t *dtl = S + kNa.tl, *dcl = S + kNa.cl, *dqc = S + kNa.qc, *dmid = S + kNa.mid, *dout = S + kNa.out;
  Job *jobs = reinterpret_cast<Job *>(S + kNa.jobs);
  uint32_t *chunks = reinterpret_cast<uint32_t *>(S + kNa.chunks), *fail = reinterpret_cast<uint32_t *>(S + kNa.fail);
  uint64_t *offs = reinterpret_cast<uint64_t *>(S + kNa.offs), *sums = reinterpret_cast<uint64_t *>(S + kNa.sums);
  uint64_t *total = reinterpret_cast<uint64_t *>(S + kNa.total);
  unsigned long long *evald = reinterpret_cast<unsigned long long *>(S + kNa.evald);
  const uint32_t nn = (uint32_t)n, pm = kPMask & ~dead_flags;
  HIPCHK(h, hipMemcpyAsync(dtw, t_words, 16 * n, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(dcw, c_words, 16 * n, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(dtl, t_len, n, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(dcl, c_len, n, hipMemcpyHostToDevice, h->stream));
  const bool vec = (stage_mask & (SX_STAGE_QC | SX_STAGE_MID)) != 0, exh = (stage_mask & SX_STAGE_EXH) != 0;
  if (vec) {
    // k_pair_vectors runs each sequence for its own length only
    hipLaunchKernelGGL(k_pair_vectors, dim3(blocks_for(n, 4)), dim3(256), 0, h->stream, dtw, dtl, dcw, dcl, n, pm,
                       dqc, dmid, (uint8_t *)nullptr);
    HIPCHK(h, hipGetLastError());
  }
  if (exh) {
    // the array keeps the last call's contents: the evaluated-state shards
    // start at zero (a fill kernel); the plan kernel writes every chunk count
    // and clears every fail flag
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(2 * kEvalShards, 256)), dim3(256), 0, h->stream,
                       reinterpret_cast<uint32_t *>(evald), (uint64_t)(2 * kEvalShards), 0u);
    hipLaunchKernelGGL(k_pair_jobs, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, dtw, dtl, dcw, dcl, nn, pm,
                       jobs);
    hipLaunchKernelGGL(k_exh_plan, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, jobs, nn, (int)exh_mode,
                       (const PlanEntry *)nullptr, pm, chunks, fail, (uint32_t *)nullptr, (uint32_t *)nullptr);
    const uint64_t nb = (n + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_tile_sums, dim3((unsigned)nb), dim3(256), 0, h->stream, chunks, n, sums, 0);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, h->stream, sums, nb, total);
    hipLaunchKernelGGL(k_tile_scan, dim3((unsigned)nb), dim3(256), 0, h->stream, chunks, n, (const uint64_t *)sums,
                       offs);
    HIPCHK(h, hipGetLastError());
    uint64_t tot = 0;
    HIPCHK(h, hipMemcpyAsync(&tot, total, sizeof(tot), hipMemcpyDeviceToHost, h->stream));
    SYNC_WD("exhaustive plan");
    // no chunk -> job map (its size depends on the sweeps): each wave finds
    // its job by searching the chunk offsets
    const uint64_t slice = 1ull << 20;
    for (uint64_t c0 = 0; c0 < tot; c0 += slice) {
      int rc = launch_exh_run(h, jobs, nn, offs, nullptr, c0, std::min<uint64_t>(slice, tot - c0), fail, evald);
      if (rc) return rc;
    }
  }
  hipLaunchKernelGGL(k_pair_bits, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, vec ? dqc : nullptr,
                     vec ? dmid : nullptr, exh ? fail : nullptr, nn, stage_mask, dout);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipMemcpyAsync(out, dout, n, hipMemcpyDeviceToHost, h->stream));
  SYNC_WD("device sync");
  return SX_OK;
}

extern "C" int sx_flag_diff(sx_handle *h, const uint32_t *t_words, const uint8_t *t_len, const uint32_t *c_words,
                            const uint8_t *c_len, uint64_t n, uint8_t *out) {
  if (!h || (n && (!t_words || !t_len || !c_words || !c_len || !out))) return SX_EINVAL;
  if (n == 0) return SX_OK;
  for (uint64_t i = 0; i < n; i++) {
    if (t_len[i] > 4 || c_len[i] > 4) return set_err(h, SX_EINVAL, "sequence longer than 4");
    for (int k = 0; k < t_len[i]; k++)
      if ((t_words[i * 4 + k] & 0xFFFFu) >= OP_COUNT) return set_err(h, SX_EOPCODE, "opcode out of range");
    for (int k = 0; k < c_len[i]; k++)
      if ((c_words[i * 4 + k] & 0xFFFFu) >= OP_COUNT) return set_err(h, SX_EOPCODE, "opcode out of range");
  }
  HIPCHK(h, hipSetDevice(h->device));
  uint32_t *dtw, *dcw;
  uint8_t *dtl, *dcl, *o3;
  ENSURE(B_P_TW, uint32_t, 4 * n, dtw);
  ENSURE(B_P_CW, uint32_t, 4 * n, dcw);
  ENSURE(B_P_TL, uint8_t, n, dtl);
  ENSURE(B_P_CL, uint8_t, n, dcl);
  ENSURE(B_P_OUT3, uint8_t, n, o3);
  HIPCHK(h, hipMemcpyAsync(dtw, t_words, 16 * n, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(dcw, c_words, 16 * n, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(dtl, t_len, n, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(dcl, c_len, n, hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(k_pair_vectors, dim3(blocks_for(n, 4)), dim3(256), 0, h->stream, dtw, dtl, dcw, dcl, n,
                     (uint32_t)kPMask, (uint8_t *)nullptr, (uint8_t *)nullptr, o3);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipMemcpyAsync(out, o3, n, hipMemcpyDeviceToHost, h->stream));
  SYNC_WD("device sync");
  return SX_OK;
}

extern "C" int sx_exec_states(sx_handle *h, uint8_t *states, const uint32_t *words, const uint8_t *len, uint64_t n) {
  if (!h || (n && (!states || !words || !len))) return SX_EINVAL;
  if (n == 0) return SX_OK;
  for (uint64_t i = 0; i < n; i++)
    if (len[i] > 4) return set_err(h, SX_EINVAL, "sequence longer than 4");
  HIPCHK(h, hipSetDevice(h->device));
  uint8_t *ds, *dl;
  uint32_t *dw, *bad;
  ENSURE(B_STATES, uint8_t, 8 * n, ds);
  ENSURE(B_P_TW, uint32_t, 4 * n, dw);
  ENSURE(B_P_TL, uint8_t, n, dl);
  ENSURE(B_BADOP, uint32_t, 1, bad);
  HIPCHK(h, hipMemcpyAsync(ds, states, 8 * n, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(dw, words, 16 * n, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(dl, len, n, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemsetAsync(bad, 0, 4, h->stream));
  hipLaunchKernelGGL(k_exec_states, dim3(blocks_for(n, 256)), dim3(256), 0, h->stream, ds, dw, dl, n, bad);
  HIPCHK(h, hipGetLastError());
  uint32_t b = 0;
  HIPCHK(h, hipMemcpyAsync(states, ds, 8 * n, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(&b, bad, 4, hipMemcpyDeviceToHost, h->stream));
  SYNC_WD("device sync");
  if (b) return set_err(h, SX_EOPCODE, "opcode out of range");
  return SX_OK;
}

extern "C" int sx_exh_plan(const uint32_t *t, int tn, const uint32_t *c, int cn, uint32_t dead_flags,
                           int32_t exh_mode, uint8_t sets[8], int32_t *n_sets, uint64_t *states) {
  if (!t || !c || tn < 0 || tn > 4 || cn < 0 || cn > 4 || !sets || !n_sets) return SX_EINVAL;
  for (int i = 0; i < tn; i++)
    if ((t[i] & 0xFFFFu) >= OP_COUNT) return SX_EOPCODE;
  for (int i = 0; i < cn; i++)
    if ((c[i] & 0xFFFFu) >= OP_COUNT) return SX_EOPCODE;
  Plan pl = make_plan(t, tn, c, cn, kPMask & ~dead_flags, exh_mode, h_reads, h_dep);
  uint64_t tot = 0;
  for (int i = 0; i < 8; i++) sets[i] = i < pl.nsets ? pl.sets[i] : 0;
  for (int i = 0; i < pl.nsets; i++) tot += 1ull << pl.logsz[i];
  *n_sets = pl.nsets;
  if (states) *states = tot;
  return SX_OK;
}

extern "C" int sx_dep_table(uint16_t out[58 * 16]) {
  if (!out) return SX_EINVAL;
  memcpy(out, h_dep.d, sizeof(h_dep.d));
  return SX_OK;
}

// ----------------------------------------------------- internal (C++) entry --

hipStream_t sx_internal_stream(sx_handle *h) { return h->stream; }

int sx_internal_error(sx_handle *h, int code, const char *msg) { return set_err(h, code, "%s", msg); }

// Complete (projected) ExhaustiveCheck of n pairs of sequences of up to 10
// words, stride 10; ok[i] = 1 when pair i passes.  Used by the STOKE driver.
int sx_verify_seqs(sx_handle *h, const uint32_t *t_words, const uint8_t *t_len, const uint32_t *c_words,
                   const uint8_t *c_len, uint64_t n, uint32_t dead_flags, uint8_t *ok) {
  if (h->poisoned) return set_err(h, SX_EHIP, "handle unusable after a device timeout");
  const uint32_t pm = kPMask & ~dead_flags;
  const uint64_t batch = 1u << 20;
  for (uint64_t b0 = 0; b0 < n; b0 += batch) {
    uint32_t nb = (uint32_t)std::min<uint64_t>(batch, n - b0);
    std::vector<Job> jobs(nb);
    for (uint32_t i = 0; i < nb; i++) {
      Job &j = jobs[i];
      memset(&j, 0, sizeof(j));
      for (int k = 0; k < kJobWords; k++) {
        j.tw[k] = k < t_len[b0 + i] ? t_words[(b0 + i) * kJobWords + k] : 0u;
        j.cw[k] = k < c_len[b0 + i] ? c_words[(b0 + i) * kJobWords + k] : 0u;
      }
      j.tl = t_len[b0 + i];
      j.cl = c_len[b0 + i];
      j.pass = 1;
      j.pmask = (uint8_t)pm;
      j.target = i;
    }
    Job *dj;
    ENSURE(B_JOBS, Job, nb, dj);
    HIPCHK(h, hipMemcpyAsync(dj, jobs.data(), nb * sizeof(Job), hipMemcpyHostToDevice, h->stream));
    uint64_t chunks = 0;
    int rc = run_exhaustive(h, nb, SX_EXH_PROJECTED, &chunks);
    if (rc) return rc;
    std::vector<uint32_t> f(nb);
    HIPCHK(h, hipMemcpyAsync(f.data(), h->bufs[B_FAIL].p, nb * 4, hipMemcpyDeviceToHost, h->stream));
    SYNC_WD("verify");
    for (uint32_t i = 0; i < nb; i++) ok[b0 + i] = f[i] ? 0 : 1;
  }
  return SX_OK;
}
