// sx_isa.h -- 6502 instruction model shared by the HIP kernels and the host
// side of lib6502core.
//
// Semantics follow pkg/cpu/exec.go:8-327 of oisee/6502-optimizer (D = 0
// binary mode, a two-deep virtual stack in S0/S1).  The state lives in eight
// 32-bit registers holding one byte each, so a step is a handful of integer
// VALU ops; callers pass a wave-uniform opcode, which turns the dispatch
// below into scalar branches.
#pragma once

#include <stdint.h>

#include <initializer_list>

#ifdef __HIPCC__
#define SX_HD __host__ __device__ __forceinline__
#else
#define SX_HD inline
#endif

namespace sx {

// Opcode ids: pkg/inst/instruction.go:42-122 (iota order).
enum Op : uint32_t {
  TAX = 0, TAY, TXA, TYA, TSX, TXS, INX, INY, DEX, DEY,
  ASL_A, LSR_A, ROL_A, ROR_A, CLC, SEC, CLD, SED, CLI, SEI, CLV,
  PHA, PLA, PHP, PLP, NOP,
  LDA_IMM, LDX_IMM, LDY_IMM, ADC_IMM, SBC_IMM, AND_IMM, ORA_IMM, EOR_IMM,
  CMP_IMM, CPX_IMM, CPY_IMM,
  LDA_M, LDX_M, LDY_M, STA_M, STX_M, STY_M, ADC_M, SBC_M, AND_M, ORA_M, EOR_M,
  CMP_M, CPX_M, CPY_M, INC_M, DEC_M, ASL_M, LSR_M, ROL_M, ROR_M, BIT_M,
  OP_COUNT
};

constexpr int kNumInsts = 2863;     // 47 + 11 * 256
constexpr int kNumNonImm = 47;
constexpr uint32_t kPMask = 0xCF;   // pkg/cpu/state.go:20

// Register byte positions in the 8-byte state (pkg/cpu/state.go:11-16).
enum Reg : int { RA = 0, RX, RY, RP, RS, RM, RS0, RS1 };

// Go register masks (pkg/search/verifier.go:134-143), used for opReads /
// opWrites and therefore for the exhaustive domain and the pruner.
enum GoReg : uint32_t { GA = 1, GP = 2, GX = 4, GY = 8, GS = 16, GM = 32, GS0 = 64, GS1 = 128 };

SX_HD bool has_imm(uint32_t op) { return op >= LDA_IMM && op <= CPY_IMM; }

// Go instruction order (pkg/search/enumerator.go:14-38): the 47 non-immediate
// opcodes in id order, then the 11 immediate opcodes x 256.  Index -> word.
SX_HD uint32_t go_inst(uint32_t i) {
  if (i < 26) return i;                       // TAX .. NOP
  if (i < kNumNonImm) return i - 26 + LDA_M;  // LDA M .. BIT M
  uint32_t k = i - kNumNonImm;
  return (LDA_IMM + (k >> 8)) | ((k & 0xFFu) << 16);
}

// Byte sizes: model 0 = Go (pkg/inst/catalog.go:50-56, the 21 M ops count one
// byte), model 1 = the CUDA tool (cuda/6502_search_v2.cu:437-440).
SX_HD int byte_size(uint32_t op, int model) {
  if (model == 1) return op <= NOP ? 1 : 2;
  return has_imm(op) ? 2 : 1;
}

// pkg/inst/catalog.go:108-200
SX_HD int cycles(uint32_t op) {
  if (op == PHA || op == PHP) return 3;
  if (op == PLA || op == PLP) return 4;
  if (op >= INC_M && op <= ROR_M) return 5;
  if (op >= LDA_M && op <= BIT_M) return 3;
  return 2;
}

// pkg/search/verifier.go:154-259 as a table indexed by opcode.
#define SX_READS_TABLE                                                          \
  {GA, GA, GX, GY, GS, GX, GX, GY, GX, GY,                                      \
   GA, GA, GA | GP, GA | GP, 0, 0, 0, 0, 0, 0, 0,                               \
   GA | GS0 | GS1, GS0, GP | GS0 | GS1, GS0, 0,                                 \
   0, 0, 0, GA | GP, GA | GP, GA, GA, GA, GA, GX, GY,                           \
   GM, GM, GM, GA, GX, GY, GA | GM | GP, GA | GM | GP, GA | GM, GA | GM, GA | GM, \
   GA | GM, GX | GM, GY | GM, GM, GM, GM, GM, GM | GP, GM | GP, GA | GM}

// pkg/search/pruner.go:76-163 as a table indexed by opcode.
#define SX_WRITES_TABLE                                                         \
  {GX | GP, GY | GP, GA | GP, GA | GP, GX | GP, GS, GX | GP, GY | GP, GX | GP,   \
   GY | GP, GA | GP, GA | GP, GA | GP, GA | GP, GP, GP, GP, GP, GP, GP, GP,      \
   GS | GS0 | GS1, GA | GP | GS | GS0, GS | GS0 | GS1, GP | GS | GS0, 0,        \
   GA | GP, GX | GP, GY | GP, GA | GP, GA | GP, GA | GP, GA | GP, GA | GP, GP,   \
   GP, GP, GA | GP, GX | GP, GY | GP, GM, GM, GM, GA | GP, GA | GP, GA | GP,     \
   GA | GP, GA | GP, GP, GP, GP, GM | GP, GM | GP, GM | GP, GM | GP, GM | GP,    \
   GM | GP, GP}

// ShouldPrune, pkg/search/pruner.go:6-181 (rd / wr: the two tables above).
SX_HD bool should_prune(const uint32_t *seq, int n, const uint8_t *rd, const uint8_t *wr) {
  for (int i = 0; i < n; i++) {
    if ((seq[i] & 0xFFFFu) == NOP) return true;
    if (i + 1 < n) {
      uint32_t w = wr[seq[i] & 0xFFFFu];
      if (w) {
        uint32_t r2 = rd[seq[i + 1] & 0xFFFFu], w2 = wr[seq[i + 1] & 0xFFFFu];
        if (w & w2 & ~(uint32_t)(GP | GS | GS0 | GS1) & ~r2) return true;
      }
    }
  }
  int depth = 0;
  for (int i = 0; i < n; i++) {
    uint32_t op = seq[i] & 0xFFFFu;
    if (op == PHA || op == PHP) {
      if (++depth > 2) return true;
    } else if (op == PLA || op == PLP) {
      if (depth <= 0) return true;
      depth--;
    }
  }
  for (int i = 0; i + 1 < n; i++) {
    uint32_t a = seq[i] & 0xFFFFu, b = seq[i + 1] & 0xFFFFu;
    uint32_t aR = rd[a], aW = wr[a], bR = rd[b], bW = wr[b];
    bool indep = !((aW & bR) || (aR & bW) || (aW & bW));
    uint32_t ka = (a << 16) | (seq[i] >> 16), kb = (b << 16) | (seq[i + 1] >> 16);
    if (indep && ka > kb) return true;
  }
  return false;
}

// Data dependencies per opcode at flag granularity.  Inputs and outputs are
// numbered: state bytes A=0, X=1, Y=2, S=4, M=5, S0=6, S1=7 (position 3, the
// flag byte, is split) and flag bits 8 + b for P bit b (C=8, Z=9, I=10, D=11,
// B=12, U=13, V=14, N=15).  Entry [op][o] is the set of inputs that output o
// is a function of, or kDepId when the op leaves o unchanged.  Derived from
// pkg/cpu/exec.go:8-327; the tests check it for soundness by brute force.
constexpr uint16_t kDepId = 0xFFFFu;
constexpr int kDepOut = 16;

struct DepTable {
  uint16_t d[OP_COUNT][kDepOut];
};

constexpr uint16_t DB(int r) { return (uint16_t)(1u << r); }
constexpr int FC_ = 8, FZ_ = 9, FI_ = 10, FD_ = 11, FV_ = 14, FN_ = 15;

constexpr DepTable make_dep_table() {
  DepTable t{};
  for (int op = 0; op < OP_COUNT; op++)
    for (int o = 0; o < kDepOut; o++) t.d[op][o] = kDepId;
  auto nz = [&](int op, uint16_t src) {
    t.d[op][FN_] = src;
    t.d[op][FZ_] = src;
  };
  // transfers
  t.d[TAX][RX] = DB(RA); nz(TAX, DB(RA));
  t.d[TAY][RY] = DB(RA); nz(TAY, DB(RA));
  t.d[TXA][RA] = DB(RX); nz(TXA, DB(RX));
  t.d[TYA][RA] = DB(RY); nz(TYA, DB(RY));
  t.d[TSX][RX] = DB(RS); nz(TSX, DB(RS));
  t.d[TXS][RS] = DB(RX);
  // inc/dec registers
  t.d[INX][RX] = DB(RX); nz(INX, DB(RX));
  t.d[DEX][RX] = DB(RX); nz(DEX, DB(RX));
  t.d[INY][RY] = DB(RY); nz(INY, DB(RY));
  t.d[DEY][RY] = DB(RY); nz(DEY, DB(RY));
  // shifts on A: result from A (and C for rotates); C from A; N,Z from the result
  t.d[ASL_A][RA] = DB(RA); t.d[ASL_A][FC_] = DB(RA); nz(ASL_A, DB(RA));
  t.d[LSR_A][RA] = DB(RA); t.d[LSR_A][FC_] = DB(RA); nz(LSR_A, DB(RA));
  t.d[ROL_A][RA] = DB(RA) | DB(FC_); t.d[ROL_A][FC_] = DB(RA); nz(ROL_A, DB(RA) | DB(FC_));
  t.d[ROR_A][RA] = DB(RA) | DB(FC_); t.d[ROR_A][FC_] = DB(RA); nz(ROR_A, DB(RA) | DB(FC_));
  // flag operations set a constant
  t.d[CLC][FC_] = 0; t.d[SEC][FC_] = 0; t.d[CLD][FD_] = 0; t.d[SED][FD_] = 0;
  t.d[CLI][FI_] = 0; t.d[SEI][FI_] = 0; t.d[CLV][FV_] = 0;
  // stack: push moves S0 to S1 and the value to S0; pull the reverse, S1 := 0
  t.d[PHA][RS] = DB(RS); t.d[PHA][RS0] = DB(RA); t.d[PHA][RS1] = DB(RS0);
  t.d[PHP][RS] = DB(RS); t.d[PHP][RS1] = DB(RS0);
  t.d[PHP][RS0] = DB(FC_) | DB(FZ_) | DB(FI_) | DB(FD_) | DB(FV_) | DB(FN_);  // B, U pushed as 1
  t.d[PLA][RS] = DB(RS); t.d[PLA][RA] = DB(RS0); t.d[PLA][RS0] = DB(RS1); t.d[PLA][RS1] = 0;
  nz(PLA, DB(RS0));
  t.d[PLP][RS] = DB(RS); t.d[PLP][RS0] = DB(RS1); t.d[PLP][RS1] = 0;
  for (int b = 8; b < 16; b++) t.d[PLP][b] = DB(RS0);
  // immediate loads
  t.d[LDA_IMM][RA] = 0; nz(LDA_IMM, 0);
  t.d[LDX_IMM][RX] = 0; nz(LDX_IMM, 0);
  t.d[LDY_IMM][RY] = 0; nz(LDY_IMM, 0);
  // ALU with an immediate or with M
  const uint16_t opnd_imm = 0, opnd_m = DB(RM);
  for (int k = 0; k < 2; k++) {
    const uint16_t opnd = k ? opnd_m : opnd_imm;
    const int adc = k ? ADC_M : ADC_IMM, sbc = k ? SBC_M : SBC_IMM;
    const int an = k ? AND_M : AND_IMM, ora = k ? ORA_M : ORA_IMM, eor = k ? EOR_M : EOR_IMM;
    const int cpa = k ? CMP_M : CMP_IMM, cpx = k ? CPX_M : CPX_IMM, cpy = k ? CPY_M : CPY_IMM;
    for (int op : {adc, sbc}) {
      const uint16_t src = DB(RA) | opnd | DB(FC_);
      t.d[op][RA] = src; t.d[op][FC_] = src; t.d[op][FV_] = src; nz(op, src);
    }
    for (int op : {an, ora, eor}) {
      t.d[op][RA] = DB(RA) | opnd; nz(op, DB(RA) | opnd);
    }
    t.d[cpa][FC_] = DB(RA) | opnd; nz(cpa, DB(RA) | opnd);
    t.d[cpx][FC_] = DB(RX) | opnd; nz(cpx, DB(RX) | opnd);
    t.d[cpy][FC_] = DB(RY) | opnd; nz(cpy, DB(RY) | opnd);
  }
  // loads and stores with M
  t.d[LDA_M][RA] = DB(RM); nz(LDA_M, DB(RM));
  t.d[LDX_M][RX] = DB(RM); nz(LDX_M, DB(RM));
  t.d[LDY_M][RY] = DB(RM); nz(LDY_M, DB(RM));
  t.d[STA_M][RM] = DB(RA);
  t.d[STX_M][RM] = DB(RX);
  t.d[STY_M][RM] = DB(RY);
  // read-modify-write on M
  t.d[INC_M][RM] = DB(RM); nz(INC_M, DB(RM));
  t.d[DEC_M][RM] = DB(RM); nz(DEC_M, DB(RM));
  t.d[ASL_M][RM] = DB(RM); t.d[ASL_M][FC_] = DB(RM); nz(ASL_M, DB(RM));
  t.d[LSR_M][RM] = DB(RM); t.d[LSR_M][FC_] = DB(RM); nz(LSR_M, DB(RM));
  t.d[ROL_M][RM] = DB(RM) | DB(FC_); t.d[ROL_M][FC_] = DB(RM); nz(ROL_M, DB(RM) | DB(FC_));
  t.d[ROR_M][RM] = DB(RM) | DB(FC_); t.d[ROR_M][FC_] = DB(RM); nz(ROR_M, DB(RM) | DB(FC_));
  // BIT: N and V from M, Z from A & M
  t.d[BIT_M][FN_] = DB(RM); t.d[BIT_M][FV_] = DB(RM); t.d[BIT_M][FZ_] = DB(RA) | DB(RM);
  return t;
}

// QuickCheck and MidCheck vectors, pkg/search/verifier.go:18-67 (A X Y P S M
// S0 S1).  The first 8 Mid vectors are the QuickCheck vectors.
#define SX_MID_VECTORS                                                                   \
  {{0x00, 0x00, 0x00, 0x00, 0xFD, 0x42, 0x00, 0x00}, {0xFF, 0xFF, 0xFF, 0xFF, 0xFF, 0xBD, 0xFF, 0xFF}, \
   {0x01, 0x02, 0x03, 0x00, 0xFD, 0x13, 0x10, 0x20}, {0x80, 0x40, 0x20, 0x01, 0xFC, 0x7E, 0x30, 0x40}, \
   {0x55, 0xAA, 0x55, 0x00, 0xFE, 0xA5, 0x55, 0xAA}, {0xAA, 0x55, 0xAA, 0x01, 0xFD, 0x5A, 0xAA, 0x55}, \
   {0x0F, 0xF0, 0x0F, 0x00, 0xFB, 0xE1, 0x0F, 0xF0}, {0x7F, 0x80, 0x7F, 0x01, 0xFD, 0x33, 0x80, 0x7F}, \
   {0x02, 0x01, 0x04, 0x00, 0xFD, 0x91, 0x11, 0x22}, {0x04, 0x08, 0x10, 0x01, 0xFD, 0x48, 0x33, 0x44}, \
   {0x08, 0x10, 0x20, 0x00, 0xFD, 0x24, 0x55, 0x66}, {0x10, 0x20, 0x40, 0x01, 0xFD, 0xD7, 0x77, 0x88}, \
   {0x20, 0x40, 0x80, 0x00, 0xFD, 0x6B, 0x99, 0xBB}, {0x40, 0x80, 0x01, 0x01, 0xFD, 0xB2, 0xCC, 0xDD}, \
   {0x03, 0x05, 0x09, 0x00, 0xFD, 0x0F, 0xEE, 0x11}, {0x05, 0x0A, 0x14, 0x01, 0xFD, 0xF0, 0x22, 0x33}, \
   {0xFE, 0x7F, 0x80, 0x00, 0xFD, 0x01, 0x44, 0x55}, {0x81, 0xFE, 0x01, 0x01, 0xFD, 0xFE, 0x66, 0x77}, \
   {0x7E, 0x81, 0x42, 0x00, 0xFD, 0x80, 0x88, 0x99}, {0xBF, 0x40, 0xBF, 0x01, 0xFD, 0x40, 0xAA, 0xBB}, \
   {0xC0, 0x3F, 0xC0, 0x00, 0xFD, 0x1F, 0xCC, 0xDD}, {0xE0, 0x1F, 0xE0, 0x01, 0xFD, 0xEF, 0xEE, 0x00}, \
   {0xF0, 0x0F, 0xF0, 0x00, 0xFD, 0x55, 0x11, 0x22}, {0x0E, 0xE0, 0x0E, 0x01, 0xFD, 0xAA, 0x33, 0x44}, \
   {0x11, 0x22, 0x33, 0x00, 0xFE, 0xC3, 0x00, 0x00}, {0x22, 0x33, 0x44, 0x01, 0xFF, 0x3C, 0x55, 0x66}, \
   {0x44, 0x55, 0x66, 0x00, 0xFC, 0x87, 0x77, 0x88}, {0x88, 0x99, 0xAA, 0x01, 0xFB, 0x69, 0x99, 0xAA}, \
   {0x33, 0x44, 0x55, 0x00, 0xFD, 0xAE, 0xBB, 0xCC}, {0xCC, 0xDD, 0xEE, 0x01, 0xFD, 0x51, 0xDD, 0xEE}, \
   {0x66, 0x77, 0x88, 0x00, 0xFD, 0xDA, 0x12, 0x34}, {0x99, 0xAA, 0xBB, 0x01, 0xFD, 0x25, 0x56, 0x78}}

// The 8-byte machine state, one byte per 32-bit register.
struct State {
  uint32_t a, x, y, p, s, m, s0, s1;
};

SX_HD uint32_t nz_flags(uint32_t v) { return (v & 0x80u) | (v == 0 ? 0x02u : 0u); }
SX_HD void set_nz(State &st, uint32_t v) { st.p = (st.p & ~0x82u) | nz_flags(v); }

// ADC / SBC with D = 0 (exec.go:202-242).
SX_HD void adc(State &st, uint32_t m) {
  uint32_t r = st.a + m + (st.p & 1u);
  uint32_t r8 = r & 0xFFu;
  uint32_t v = (st.a ^ r8) & (m ^ r8) & 0x80u;
  st.p = (st.p & ~0xC3u) | nz_flags(r8) | (r >> 8) | (v >> 1);
  st.a = r8;
}
SX_HD void sbc(State &st, uint32_t m) {
  // a - m - borrow; carry out = no borrow = (a + (m ^ 0xFF) + c) > 0xFF
  uint32_t r = st.a + (m ^ 0xFFu) + (st.p & 1u);
  uint32_t r8 = r & 0xFFu;
  uint32_t v = (st.a ^ r8) & (st.a ^ m) & 0x80u;
  st.p = (st.p & ~0xC3u) | nz_flags(r8) | (r >> 8) | (v >> 1);
  st.a = r8;
}
SX_HD void cmp(State &st, uint32_t reg, uint32_t m) {
  uint32_t r = (reg - m) & 0xFFu;
  st.p = (st.p & ~0x83u) | nz_flags(r) | (reg >= m ? 1u : 0u);
}
SX_HD uint32_t asl(State &st, uint32_t v) {
  uint32_t r = (v << 1) & 0xFFu;
  st.p = (st.p & ~0x83u) | nz_flags(r) | (v >> 7);
  return r;
}
SX_HD uint32_t lsr(State &st, uint32_t v) {
  uint32_t r = v >> 1;
  st.p = (st.p & ~0x83u) | nz_flags(r) | (v & 1u);
  return r;
}
SX_HD uint32_t rol(State &st, uint32_t v) {
  uint32_t r = ((v << 1) | (st.p & 1u)) & 0xFFu;
  st.p = (st.p & ~0x83u) | nz_flags(r) | (v >> 7);
  return r;
}
SX_HD uint32_t ror(State &st, uint32_t v) {
  uint32_t r = (v >> 1) | ((st.p & 1u) << 7);
  st.p = (st.p & ~0x83u) | nz_flags(r) | (v & 1u);
  return r;
}

// One instruction step over K states, pkg/cpu/exec.go:8-189: the opcode is
// dispatched once and its body applied to every state.  Returns false on an
// unknown opcode (where Go panics); the states are then left unchanged.
#define SX_EACH(...)                       \
  {                                        \
    for (int j_ = 0; j_ < K; j_++) {       \
      State &s = st[j_];                   \
      __VA_ARGS__;                         \
    }                                      \
  }                                        \
  break;
template <int K>
SX_HD bool step_many(State (&st)[K], uint32_t op, uint32_t imm) {
  imm &= 0xFFu;
  switch (op) {
  case TAX: SX_EACH(s.x = s.a; set_nz(s, s.x);)
  case TAY: SX_EACH(s.y = s.a; set_nz(s, s.y);)
  case TXA: SX_EACH(s.a = s.x; set_nz(s, s.a);)
  case TYA: SX_EACH(s.a = s.y; set_nz(s, s.a);)
  case TSX: SX_EACH(s.x = s.s; set_nz(s, s.x);)
  case TXS: SX_EACH(s.s = s.x;)
  case INX: SX_EACH(s.x = (s.x + 1) & 0xFFu; set_nz(s, s.x);)
  case INY: SX_EACH(s.y = (s.y + 1) & 0xFFu; set_nz(s, s.y);)
  case DEX: SX_EACH(s.x = (s.x - 1) & 0xFFu; set_nz(s, s.x);)
  case DEY: SX_EACH(s.y = (s.y - 1) & 0xFFu; set_nz(s, s.y);)
  case ASL_A: SX_EACH(s.a = asl(s, s.a);)
  case LSR_A: SX_EACH(s.a = lsr(s, s.a);)
  case ROL_A: SX_EACH(s.a = rol(s, s.a);)
  case ROR_A: SX_EACH(s.a = ror(s, s.a);)
  case CLC: SX_EACH(s.p &= ~0x01u;)
  case SEC: SX_EACH(s.p |= 0x01u;)
  case CLD: SX_EACH(s.p &= ~0x08u;)
  case SED: SX_EACH(s.p |= 0x08u;)
  case CLI: SX_EACH(s.p &= ~0x04u;)
  case SEI: SX_EACH(s.p |= 0x04u;)
  case CLV: SX_EACH(s.p &= ~0x40u;)
  case PHA: SX_EACH(s.s1 = s.s0; s.s0 = s.a; s.s = (s.s - 1) & 0xFFu;)
  case PLA: SX_EACH(s.a = s.s0; s.s0 = s.s1; s.s1 = 0; s.s = (s.s + 1) & 0xFFu; set_nz(s, s.a);)
  case PHP: SX_EACH(s.s1 = s.s0; s.s0 = s.p | 0x30u; s.s = (s.s - 1) & 0xFFu;)
  case PLP: SX_EACH(s.p = s.s0; s.s0 = s.s1; s.s1 = 0; s.s = (s.s + 1) & 0xFFu;)
  case NOP: SX_EACH()
  case LDA_IMM: SX_EACH(s.a = imm; set_nz(s, s.a);)
  case LDX_IMM: SX_EACH(s.x = imm; set_nz(s, s.x);)
  case LDY_IMM: SX_EACH(s.y = imm; set_nz(s, s.y);)
  case ADC_IMM: SX_EACH(adc(s, imm);)
  case SBC_IMM: SX_EACH(sbc(s, imm);)
  case AND_IMM: SX_EACH(s.a &= imm; set_nz(s, s.a);)
  case ORA_IMM: SX_EACH(s.a |= imm; set_nz(s, s.a);)
  case EOR_IMM: SX_EACH(s.a ^= imm; set_nz(s, s.a);)
  case CMP_IMM: SX_EACH(cmp(s, s.a, imm);)
  case CPX_IMM: SX_EACH(cmp(s, s.x, imm);)
  case CPY_IMM: SX_EACH(cmp(s, s.y, imm);)
  case LDA_M: SX_EACH(s.a = s.m; set_nz(s, s.a);)
  case LDX_M: SX_EACH(s.x = s.m; set_nz(s, s.x);)
  case LDY_M: SX_EACH(s.y = s.m; set_nz(s, s.y);)
  case STA_M: SX_EACH(s.m = s.a;)
  case STX_M: SX_EACH(s.m = s.x;)
  case STY_M: SX_EACH(s.m = s.y;)
  case ADC_M: SX_EACH(adc(s, s.m);)
  case SBC_M: SX_EACH(sbc(s, s.m);)
  case AND_M: SX_EACH(s.a &= s.m; set_nz(s, s.a);)
  case ORA_M: SX_EACH(s.a |= s.m; set_nz(s, s.a);)
  case EOR_M: SX_EACH(s.a ^= s.m; set_nz(s, s.a);)
  case CMP_M: SX_EACH(cmp(s, s.a, s.m);)
  case CPX_M: SX_EACH(cmp(s, s.x, s.m);)
  case CPY_M: SX_EACH(cmp(s, s.y, s.m);)
  case INC_M: SX_EACH(s.m = (s.m + 1) & 0xFFu; set_nz(s, s.m);)
  case DEC_M: SX_EACH(s.m = (s.m - 1) & 0xFFu; set_nz(s, s.m);)
  case ASL_M: SX_EACH(s.m = asl(s, s.m);)
  case LSR_M: SX_EACH(s.m = lsr(s, s.m);)
  case ROL_M: SX_EACH(s.m = rol(s, s.m);)
  case ROR_M: SX_EACH(s.m = ror(s, s.m);)
  case BIT_M: SX_EACH(s.p = (s.p & ~0xC2u) | (s.m & 0xC0u) | ((s.a & s.m) == 0 ? 0x02u : 0u);)
  default: return false;
  }
  return true;
}
#undef SX_EACH

// One instruction step on one state.
SX_HD bool step(State &st, uint32_t op, uint32_t imm) {
  return step_many<1>(*reinterpret_cast<State(*)[1]>(&st), op, imm);
}

// Pack / unpack to two little-endian words: lo = A|X<<8|Y<<16|P<<24,
// hi = S|M<<8|S0<<16|S1<<24 (the state's byte order).
SX_HD uint32_t pack_lo(const State &s) { return s.a | (s.x << 8) | (s.y << 16) | (s.p << 24); }
SX_HD uint32_t pack_hi(const State &s) { return s.s | (s.m << 8) | (s.s0 << 16) | (s.s1 << 24); }
SX_HD State unpack(uint32_t lo, uint32_t hi) {
  State s;
  s.a = lo & 0xFFu; s.x = (lo >> 8) & 0xFFu; s.y = (lo >> 16) & 0xFFu; s.p = lo >> 24;
  s.s = hi & 0xFFu; s.m = (hi >> 8) & 0xFFu; s.s0 = (hi >> 16) & 0xFFu; s.s1 = hi >> 24;
  return s;
}

}  // namespace sx
