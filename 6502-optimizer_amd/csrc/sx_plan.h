// sx_plan.h -- the exhaustive-check sweep plan of one (target, candidate)
// pair, shared by the device kernels and the host.
//
// The domain is that of pkg/search/verifier.go:262-415: A over 0..255, P over
// carry 0/1 (other flag bits 0), S = 0xFD unless S is read, and X, Y, M, S,
// S0, S1 over 0..255 when either sequence reads them (opReads), 0 otherwise.
//
// SX_EXH_FULL sweeps that whole domain: 2^(9 + 8k) states for k extra
// registers.  SX_EXH_PROJECTED splits it per output byte: output byte r of
// either sequence is a function of the input registers in its dependency set
// (composed from the per-opcode table in sx_isa.h), so "T and C agree on byte
// r for every state of the domain" holds iff it holds on the sub-sweep that
// varies only the registers of that set that the domain varies (the others
// stay at their single domain value).  Each sub-sweep compares the whole
// state, so a mismatch anywhere is a real counterexample; the pair passes iff
// every sub-sweep passes.  Sub-sweeps whose register set is contained in
// another are dropped (the larger sweep visits all their states).
#pragma once

#include "sx_isa.h"

namespace sx {

struct Plan {
  uint8_t nsets;
  uint8_t sets[8];   // bit r = state byte r varies
  uint8_t logsz[8];  // log2 of the sub-sweep size
};

SX_HD int set_log_size(uint32_t set) {
  int n = 0;
  for (int r = 0; r < 8; r++)
    if (set & (1u << r)) n += (r == RP) ? 1 : 8;
  return n;
}

// Longest sequence a plan is made for (the sweep jobs hold up to 10 words).
constexpr int kPlanMaxSeq = 10;

// The reads table and the dependency table are passed in so that host and
// device use their own copies.  Every array index below is a compile-time
// constant (loops over sequence positions are unrolled and guarded, the kept
// sub-sweeps are packed into 64-bit words), so on the device the plan stays
// in registers.
SX_HD Plan make_plan(const uint32_t *t, int tn, const uint32_t *c, int cn, uint32_t pmask,
                     int mode, const uint8_t *reads_tab, const DepTable &dep) {
  uint32_t reads = 0;
#pragma unroll
  for (int i = 0; i < kPlanMaxSeq; i++) {
    if (i < tn) reads |= reads_tab[t[i] & 0xFFFFu];
    if (i < cn) reads |= reads_tab[c[i] & 0xFFFFu];
  }
  uint32_t vary = (1u << RA) | (1u << RP);
  if (reads & GX) vary |= 1u << RX;
  if (reads & GY) vary |= 1u << RY;
  if (reads & GM) vary |= 1u << RM;
  if (reads & GS) vary |= 1u << RS;
  if (reads & GS0) vary |= 1u << RS0;
  if (reads & GS1) vary |= 1u << RS1;
  Plan pl;
  pl.nsets = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) pl.sets[i] = 0, pl.logsz[i] = 0;
  if (mode == 1) {
    pl.nsets = 1;
    pl.sets[0] = (uint8_t)vary;
    pl.logsz[0] = (uint8_t)set_log_size(vary);
    return pl;
  }
  // dependency sets in the 16-input numbering of DepTable: only the carry
  // among the flag bits varies in the domain (the others are 0)
  const uint32_t vary16 = (vary & ~(1u << RP)) | (1u << 8);
  uint32_t dt[kDepOut], dc[kDepOut];
#pragma unroll
  for (int o = 0; o < kDepOut; o++) dt[o] = dc[o] = 1u << o;
  // compose one instruction's table into d: output o now depends on the
  // inputs of every q its row names (or keeps d[o] when unchanged)
  auto compose = [&](uint32_t (&d)[kDepOut], uint32_t word) {
    const uint16_t *row = dep.d[word & 0xFFFFu];
    uint32_t nd[kDepOut];
#pragma unroll
    for (int o = 0; o < kDepOut; o++) {
      const uint32_t r = row[o];
      uint32_t acc = 0;
#pragma unroll
      for (int q = 0; q < kDepOut; q++) acc |= (r >> q) & 1u ? d[q] : 0u;
      nd[o] = r == kDepId ? d[o] : acc;
    }
#pragma unroll
    for (int o = 0; o < kDepOut; o++) d[o] = nd[o];
  };
  // the words as shift registers, so that the loop reads only word 0
  uint32_t tw[kPlanMaxSeq], cw[kPlanMaxSeq];
#pragma unroll
  for (int i = 0; i < kPlanMaxSeq; i++) {
    tw[i] = i < tn ? t[i] : 0u;
    cw[i] = i < cn ? c[i] : 0u;
  }
  const int nmax = tn > cn ? tn : cn;
  for (int i = 0; i < nmax; i++) {
    if (i < tn) compose(dt, tw[0]);
    if (i < cn) compose(dc, cw[0]);
#pragma unroll
    for (int k = 0; k + 1 < kPlanMaxSeq; k++) {
      tw[k] = tw[k + 1];
      cw[k] = cw[k + 1];
    }
  }
  // one candidate sub-sweep per compared output (0 = none): the flag byte
  // is split into its bits, and uncompared flags are skipped
  uint32_t cand[kDepOut];
#pragma unroll
  for (int o = 0; o < kDepOut; o++) {
    const bool skip = o == RP || (o >= 8 && !((pmask & kPMask) & (1u << (o - 8))));
    const uint32_t v16 = (dt[o] | dc[o]) & vary16;
    cand[o] = skip ? 0u : ((v16 & 0xF7u) | (((v16 >> 8) & 1u) << RP)) | 0x100u;  // bit 8 = present
  }
  // keep maximal sets only (dedupe equal sets, keeping the first; drop strict
  // subsets), in output order
  uint64_t sets = 0, logs = 0;
  uint32_t n = 0;
#pragma unroll
  for (int o = 0; o < kDepOut; o++) {
    const uint32_t ci = cand[o] & 0xFFu;
    bool keep = cand[o] != 0;
#pragma unroll
    for (int p = 0; p < kDepOut; p++) {
      if (p == o) continue;
      const uint32_t cj = cand[p] & 0xFFu;
      const bool sub = cand[p] != 0 && (ci & ~cj) == 0;
      if (sub && (ci != cj || p < o)) keep = false;
    }
    if (keep && n < 8) {
      sets |= (uint64_t)ci << (8 * n);
      logs |= (uint64_t)set_log_size(ci) << (8 * n);
      n++;
    }
  }
  pl.nsets = (uint8_t)n;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    pl.sets[i] = (uint8_t)(sets >> (8 * i));
    pl.logsz[i] = (uint8_t)(logs >> (8 * i));
  }
  return pl;
}

}  // namespace sx
