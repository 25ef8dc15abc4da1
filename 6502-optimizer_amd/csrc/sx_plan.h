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

// The reads table and the dependency table are passed in so that host and
// device use their own copies.
SX_HD Plan make_plan(const uint32_t *t, int tn, const uint32_t *c, int cn, uint32_t pmask,
                     int mode, const uint8_t *reads_tab, const DepTable &dep) {
  uint32_t reads = 0;
  for (int i = 0; i < tn; i++) reads |= reads_tab[t[i] & 0xFFFFu];
  for (int i = 0; i < cn; i++) reads |= reads_tab[c[i] & 0xFFFFu];
  uint32_t vary = (1u << RA) | (1u << RP);
  if (reads & GX) vary |= 1u << RX;
  if (reads & GY) vary |= 1u << RY;
  if (reads & GM) vary |= 1u << RM;
  if (reads & GS) vary |= 1u << RS;
  if (reads & GS0) vary |= 1u << RS0;
  if (reads & GS1) vary |= 1u << RS1;
  Plan pl;
  pl.nsets = 0;
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
  for (int o = 0; o < kDepOut; o++) dt[o] = dc[o] = 1u << o;
  for (int side = 0; side < 2; side++) {
    const uint32_t *seq = side ? c : t;
    const int n = side ? cn : tn;
    uint32_t *d = side ? dc : dt;
    for (int i = 0; i < n; i++) {
      const uint16_t *row = dep.d[seq[i] & 0xFFFFu];
      uint32_t nd[kDepOut];
      for (int o = 0; o < kDepOut; o++) {
        if (row[o] == kDepId) {
          nd[o] = d[o];
        } else {
          uint32_t acc = 0;
          for (int q = 0; q < kDepOut; q++)
            if (row[o] & (1u << q)) acc |= d[q];
          nd[o] = acc;
        }
      }
      for (int o = 0; o < kDepOut; o++) d[o] = nd[o];
    }
  }
  uint32_t cand[16];
  int ncand = 0;
  for (int o = 0; o < kDepOut; o++) {
    if (o == RP) continue;                                     // the flag byte is split
    if (o >= 8 && !((pmask & kPMask) & (1u << (o - 8)))) continue;  // flag not compared
    const uint32_t v16 = (dt[o] | dc[o]) & vary16;
    cand[ncand++] = (v16 & 0xF7u) | (((v16 >> 8) & 1u) << RP);
  }
  // keep maximal sets only (dedupe equal sets, drop strict subsets)
  for (int i = 0; i < ncand; i++) {
    bool keep = true;
    for (int j = 0; j < ncand && keep; j++) {
      if (i == j) continue;
      bool sub = (cand[i] & ~cand[j]) == 0;
      if (sub && (cand[i] != cand[j] || j < i)) keep = false;
    }
    if (keep && pl.nsets < 8) {
      pl.sets[pl.nsets] = (uint8_t)cand[i];
      pl.logsz[pl.nsets] = (uint8_t)set_log_size(cand[i]);
      pl.nsets++;
    }
  }
  return pl;
}

}  // namespace sx
