// sx_pk.h -- the 6502 step of pkg/cpu/exec.go:8-189 on two states at once.
//
// Each register holds two 8-bit values in the low bytes of its 16-bit
// halves, so one packed 16-bit VALU op (v_pk_add_u16, v_pk_sub_u16, packed
// shifts) or one 32-bit bitwise op advances both states.  Every field stays in
// 0..255 per half between instructions; sums up to 511 and wrapped
// differences fit a half without crossing into the other.  The flag
// arithmetic is branch-free:
//   Z: (v - 1) >> 8 is 0xFF for v = 0 and 0 for v in 1..255, so
//      ((v - 1) >> 7) & 2 is the Z bit;
//   C of a compare: reg - m wraps to >= 0xFF01 exactly when reg < m, so
//      ~(reg - m) >> 15 is the carry (no borrow).
// The single-state step in sx_isa.h stays the reference
// semantics, and the sweep kernel compares both against the oracle in tests.
#pragma once

#include "sx_isa.h"

namespace sx {

typedef unsigned short pk16 __attribute__((ext_vector_type(2)));

SX_HD pk16 pk_splat(uint32_t c) { return (pk16)(unsigned short)c; }

struct PState {
  pk16 a, x, y, p, s, m, s0, s1;
};

SX_HD pk16 pk_nz(pk16 v) {
  return (v & pk_splat(0x80)) | (((v - pk_splat(1)) >> 7) & pk_splat(2));
}
SX_HD void pk_set_nz(PState &s, pk16 v) { s.p = (s.p & pk_splat(0x7D)) | pk_nz(v); }

SX_HD void pk_adc(PState &s, pk16 m) {
  const pk16 r = s.a + m + (s.p & pk_splat(1));
  const pk16 r8 = r & pk_splat(0xFF);
  const pk16 v = ((s.a ^ r8) & (m ^ r8) & pk_splat(0x80)) >> 1;
  s.p = (s.p & pk_splat(0x3C)) | pk_nz(r8) | (r >> 8) | v;
  s.a = r8;
}
SX_HD void pk_sbc(PState &s, pk16 m) {
  const pk16 r = s.a + (m ^ pk_splat(0xFF)) + (s.p & pk_splat(1));
  const pk16 r8 = r & pk_splat(0xFF);
  const pk16 v = ((s.a ^ r8) & (s.a ^ m) & pk_splat(0x80)) >> 1;
  s.p = (s.p & pk_splat(0x3C)) | pk_nz(r8) | (r >> 8) | v;
  s.a = r8;
}
SX_HD void pk_cmp(PState &s, pk16 reg, pk16 m) {
  const pk16 d = reg - m;
  s.p = (s.p & pk_splat(0x7C)) | pk_nz(d & pk_splat(0xFF)) | (~d >> 15);
}
SX_HD pk16 pk_asl(PState &s, pk16 v) {
  const pk16 r = (v << 1) & pk_splat(0xFF);
  s.p = (s.p & pk_splat(0x7C)) | pk_nz(r) | (v >> 7);
  return r;
}
SX_HD pk16 pk_lsr(PState &s, pk16 v) {
  const pk16 r = v >> 1;
  s.p = (s.p & pk_splat(0x7C)) | pk_nz(r) | (v & pk_splat(1));
  return r;
}
SX_HD pk16 pk_rol(PState &s, pk16 v) {
  const pk16 r = ((v << 1) | (s.p & pk_splat(1))) & pk_splat(0xFF);
  s.p = (s.p & pk_splat(0x7C)) | pk_nz(r) | (v >> 7);
  return r;
}
SX_HD pk16 pk_ror(PState &s, pk16 v) {
  const pk16 r = (v >> 1) | ((s.p & pk_splat(1)) << 7);
  s.p = (s.p & pk_splat(0x7C)) | pk_nz(r) | (v & pk_splat(1));
  return r;
}

// State registers an opcode's step below reads or writes (bit r = register
// r of enum Reg), for interpreters compiled for a subset of the registers.
constexpr uint32_t pk_touch(uint32_t op) {
  constexpr uint32_t A = 1u << RA, X = 1u << RX, Y = 1u << RY, P = 1u << RP, S = 1u << RS, M = 1u << RM,
                     S0 = 1u << RS0, S1 = 1u << RS1;
  switch (op) {
  case TAX: case TXA: return A | X | P;
  case TAY: case TYA: return A | Y | P;
  case TSX: return S | X | P;
  case TXS: return X | S;
  case INX: case DEX: return X | P;
  case INY: case DEY: return Y | P;
  case ASL_A: case LSR_A: case ROL_A: case ROR_A: return A | P;
  case CLC: case SEC: case CLD: case SED: case CLI: case SEI: case CLV: return P;
  case PHA: return A | S | S0 | S1;
  case PLA: return A | P | S | S0 | S1;
  case PHP: case PLP: return P | S | S0 | S1;
  case NOP: return 0;
  case LDA_IMM: case ADC_IMM: case SBC_IMM: case AND_IMM: case ORA_IMM: case EOR_IMM: case CMP_IMM: return A | P;
  case LDX_IMM: case CPX_IMM: return X | P;
  case LDY_IMM: case CPY_IMM: return Y | P;
  case LDA_M: case ADC_M: case SBC_M: case AND_M: case ORA_M: case EOR_M: case CMP_M: case BIT_M: return A | M | P;
  case LDX_M: case CPX_M: return X | M | P;
  case LDY_M: case CPY_M: return Y | M | P;
  case STA_M: return A | M;
  case STX_M: return X | M;
  case STY_M: return Y | M;
  default: return M | P;  // INC_M DEC_M ASL_M LSR_M ROL_M ROR_M
  }
}
constexpr uint32_t kPkAllRegs = 0xFFu;

// One instruction on K packed pairs (2K states), dispatched once; imm is the
// immediate in both halves (wave-uniform, or per lane when the lanes run
// different instances of the same opcodes).  Regs: the registers the caller
// keeps; opcodes touching others are compiled out (false, as for an unknown
// opcode), so registers outside Regs stay at their initial values.
#define SX_PK_EACH(OP, ...)                                   \
  case OP:                                                    \
    if constexpr ((pk_touch(OP) & ~Regs) == 0u) {             \
      for (int j_ = 0; j_ < K; j_++) {                        \
        PState &s = st[j_];                                   \
        __VA_ARGS__;                                          \
      }                                                       \
    } else {                                                  \
      return false;                                           \
    }                                                         \
    break;
template <int K, uint32_t Regs = kPkAllRegs>
SX_HD bool step_pk_imm(PState (&st)[K], uint32_t op, pk16 imm) {
  const pk16 one = pk_splat(1), ff = pk_splat(0xFF);
  (void)one;
  (void)ff;
  (void)imm;
  switch (op) {
  SX_PK_EACH(TAX, s.x = s.a; pk_set_nz(s, s.x))
  SX_PK_EACH(TAY, s.y = s.a; pk_set_nz(s, s.y))
  SX_PK_EACH(TXA, s.a = s.x; pk_set_nz(s, s.a))
  SX_PK_EACH(TYA, s.a = s.y; pk_set_nz(s, s.a))
  SX_PK_EACH(TSX, s.x = s.s; pk_set_nz(s, s.x))
  SX_PK_EACH(TXS, s.s = s.x)
  SX_PK_EACH(INX, s.x = (s.x + one) & ff; pk_set_nz(s, s.x))
  SX_PK_EACH(INY, s.y = (s.y + one) & ff; pk_set_nz(s, s.y))
  SX_PK_EACH(DEX, s.x = (s.x - one) & ff; pk_set_nz(s, s.x))
  SX_PK_EACH(DEY, s.y = (s.y - one) & ff; pk_set_nz(s, s.y))
  SX_PK_EACH(ASL_A, s.a = pk_asl(s, s.a))
  SX_PK_EACH(LSR_A, s.a = pk_lsr(s, s.a))
  SX_PK_EACH(ROL_A, s.a = pk_rol(s, s.a))
  SX_PK_EACH(ROR_A, s.a = pk_ror(s, s.a))
  SX_PK_EACH(CLC, s.p &= pk_splat(0xFE))
  SX_PK_EACH(SEC, s.p |= pk_splat(0x01))
  SX_PK_EACH(CLD, s.p &= pk_splat(0xF7))
  SX_PK_EACH(SED, s.p |= pk_splat(0x08))
  SX_PK_EACH(CLI, s.p &= pk_splat(0xFB))
  SX_PK_EACH(SEI, s.p |= pk_splat(0x04))
  SX_PK_EACH(CLV, s.p &= pk_splat(0xBF))
  SX_PK_EACH(PHA, s.s1 = s.s0; s.s0 = s.a; s.s = (s.s - one) & ff)
  SX_PK_EACH(PLA, s.a = s.s0; s.s0 = s.s1; s.s1 = pk_splat(0); s.s = (s.s + one) & ff; pk_set_nz(s, s.a))
  SX_PK_EACH(PHP, s.s1 = s.s0; s.s0 = s.p | pk_splat(0x30); s.s = (s.s - one) & ff)
  SX_PK_EACH(PLP, s.p = s.s0; s.s0 = s.s1; s.s1 = pk_splat(0); s.s = (s.s + one) & ff)
  case NOP: break;
  SX_PK_EACH(LDA_IMM, s.a = imm; pk_set_nz(s, s.a))
  SX_PK_EACH(LDX_IMM, s.x = imm; pk_set_nz(s, s.x))
  SX_PK_EACH(LDY_IMM, s.y = imm; pk_set_nz(s, s.y))
  SX_PK_EACH(ADC_IMM, pk_adc(s, imm))
  SX_PK_EACH(SBC_IMM, pk_sbc(s, imm))
  SX_PK_EACH(AND_IMM, s.a &= imm; pk_set_nz(s, s.a))
  SX_PK_EACH(ORA_IMM, s.a |= imm; pk_set_nz(s, s.a))
  SX_PK_EACH(EOR_IMM, s.a ^= imm; pk_set_nz(s, s.a))
  SX_PK_EACH(CMP_IMM, pk_cmp(s, s.a, imm))
  SX_PK_EACH(CPX_IMM, pk_cmp(s, s.x, imm))
  SX_PK_EACH(CPY_IMM, pk_cmp(s, s.y, imm))
  SX_PK_EACH(LDA_M, s.a = s.m; pk_set_nz(s, s.a))
  SX_PK_EACH(LDX_M, s.x = s.m; pk_set_nz(s, s.x))
  SX_PK_EACH(LDY_M, s.y = s.m; pk_set_nz(s, s.y))
  SX_PK_EACH(STA_M, s.m = s.a)
  SX_PK_EACH(STX_M, s.m = s.x)
  SX_PK_EACH(STY_M, s.m = s.y)
  SX_PK_EACH(ADC_M, pk_adc(s, s.m))
  SX_PK_EACH(SBC_M, pk_sbc(s, s.m))
  SX_PK_EACH(AND_M, s.a &= s.m; pk_set_nz(s, s.a))
  SX_PK_EACH(ORA_M, s.a |= s.m; pk_set_nz(s, s.a))
  SX_PK_EACH(EOR_M, s.a ^= s.m; pk_set_nz(s, s.a))
  SX_PK_EACH(CMP_M, pk_cmp(s, s.a, s.m))
  SX_PK_EACH(CPX_M, pk_cmp(s, s.x, s.m))
  SX_PK_EACH(CPY_M, pk_cmp(s, s.y, s.m))
  SX_PK_EACH(INC_M, s.m = (s.m + one) & ff; pk_set_nz(s, s.m))
  SX_PK_EACH(DEC_M, s.m = (s.m - one) & ff; pk_set_nz(s, s.m))
  SX_PK_EACH(ASL_M, s.m = pk_asl(s, s.m))
  SX_PK_EACH(LSR_M, s.m = pk_lsr(s, s.m))
  SX_PK_EACH(ROL_M, s.m = pk_rol(s, s.m))
  SX_PK_EACH(ROR_M, s.m = pk_ror(s, s.m))
  SX_PK_EACH(BIT_M, s.p = (s.p & pk_splat(0x3D)) | (s.m & pk_splat(0xC0)) |
                          ((((s.a & s.m) - one) >> 7) & pk_splat(2)))
  default: return false;
  }
  return true;
}
#undef SX_PK_EACH

template <int K>
SX_HD bool step_pk(PState (&st)[K], uint32_t op, uint32_t imm_word) {
  return step_pk_imm<K>(st, op, pk_splat(imm_word & 0xFFu));
}

}  // namespace sx
