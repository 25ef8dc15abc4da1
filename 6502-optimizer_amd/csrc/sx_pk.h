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

// One instruction on K packed pairs (2K states), dispatched once.
#define SX_PK_EACH(...)                   \
  {                                       \
    for (int j_ = 0; j_ < K; j_++) {      \
      PState &s = st[j_];                 \
      __VA_ARGS__;                        \
    }                                     \
  }                                       \
  break;
template <int K>
SX_HD bool step_pk(PState (&st)[K], uint32_t op, uint32_t imm_word) {
  const pk16 imm = pk_splat(imm_word & 0xFFu);
  const pk16 one = pk_splat(1), ff = pk_splat(0xFF);
  switch (op) {
  case TAX: SX_PK_EACH(s.x = s.a; pk_set_nz(s, s.x))
  case TAY: SX_PK_EACH(s.y = s.a; pk_set_nz(s, s.y))
  case TXA: SX_PK_EACH(s.a = s.x; pk_set_nz(s, s.a))
  case TYA: SX_PK_EACH(s.a = s.y; pk_set_nz(s, s.a))
  case TSX: SX_PK_EACH(s.x = s.s; pk_set_nz(s, s.x))
  case TXS: SX_PK_EACH(s.s = s.x)
  case INX: SX_PK_EACH(s.x = (s.x + one) & ff; pk_set_nz(s, s.x))
  case INY: SX_PK_EACH(s.y = (s.y + one) & ff; pk_set_nz(s, s.y))
  case DEX: SX_PK_EACH(s.x = (s.x - one) & ff; pk_set_nz(s, s.x))
  case DEY: SX_PK_EACH(s.y = (s.y - one) & ff; pk_set_nz(s, s.y))
  case ASL_A: SX_PK_EACH(s.a = pk_asl(s, s.a))
  case LSR_A: SX_PK_EACH(s.a = pk_lsr(s, s.a))
  case ROL_A: SX_PK_EACH(s.a = pk_rol(s, s.a))
  case ROR_A: SX_PK_EACH(s.a = pk_ror(s, s.a))
  case CLC: SX_PK_EACH(s.p &= pk_splat(0xFE))
  case SEC: SX_PK_EACH(s.p |= pk_splat(0x01))
  case CLD: SX_PK_EACH(s.p &= pk_splat(0xF7))
  case SED: SX_PK_EACH(s.p |= pk_splat(0x08))
  case CLI: SX_PK_EACH(s.p &= pk_splat(0xFB))
  case SEI: SX_PK_EACH(s.p |= pk_splat(0x04))
  case CLV: SX_PK_EACH(s.p &= pk_splat(0xBF))
  case PHA: SX_PK_EACH(s.s1 = s.s0; s.s0 = s.a; s.s = (s.s - one) & ff)
  case PLA: SX_PK_EACH(s.a = s.s0; s.s0 = s.s1; s.s1 = pk_splat(0); s.s = (s.s + one) & ff; pk_set_nz(s, s.a))
  case PHP: SX_PK_EACH(s.s1 = s.s0; s.s0 = s.p | pk_splat(0x30); s.s = (s.s - one) & ff)
  case PLP: SX_PK_EACH(s.p = s.s0; s.s0 = s.s1; s.s1 = pk_splat(0); s.s = (s.s + one) & ff)
  case NOP: SX_PK_EACH()
  case LDA_IMM: SX_PK_EACH(s.a = imm; pk_set_nz(s, s.a))
  case LDX_IMM: SX_PK_EACH(s.x = imm; pk_set_nz(s, s.x))
  case LDY_IMM: SX_PK_EACH(s.y = imm; pk_set_nz(s, s.y))
  case ADC_IMM: SX_PK_EACH(pk_adc(s, imm))
  case SBC_IMM: SX_PK_EACH(pk_sbc(s, imm))
  case AND_IMM: SX_PK_EACH(s.a &= imm; pk_set_nz(s, s.a))
  case ORA_IMM: SX_PK_EACH(s.a |= imm; pk_set_nz(s, s.a))
  case EOR_IMM: SX_PK_EACH(s.a ^= imm; pk_set_nz(s, s.a))
  case CMP_IMM: SX_PK_EACH(pk_cmp(s, s.a, imm))
  case CPX_IMM: SX_PK_EACH(pk_cmp(s, s.x, imm))
  case CPY_IMM: SX_PK_EACH(pk_cmp(s, s.y, imm))
  case LDA_M: SX_PK_EACH(s.a = s.m; pk_set_nz(s, s.a))
  case LDX_M: SX_PK_EACH(s.x = s.m; pk_set_nz(s, s.x))
  case LDY_M: SX_PK_EACH(s.y = s.m; pk_set_nz(s, s.y))
  case STA_M: SX_PK_EACH(s.m = s.a)
  case STX_M: SX_PK_EACH(s.m = s.x)
  case STY_M: SX_PK_EACH(s.m = s.y)
  case ADC_M: SX_PK_EACH(pk_adc(s, s.m))
  case SBC_M: SX_PK_EACH(pk_sbc(s, s.m))
  case AND_M: SX_PK_EACH(s.a &= s.m; pk_set_nz(s, s.a))
  case ORA_M: SX_PK_EACH(s.a |= s.m; pk_set_nz(s, s.a))
  case EOR_M: SX_PK_EACH(s.a ^= s.m; pk_set_nz(s, s.a))
  case CMP_M: SX_PK_EACH(pk_cmp(s, s.a, s.m))
  case CPX_M: SX_PK_EACH(pk_cmp(s, s.x, s.m))
  case CPY_M: SX_PK_EACH(pk_cmp(s, s.y, s.m))
  case INC_M: SX_PK_EACH(s.m = (s.m + one) & ff; pk_set_nz(s, s.m))
  case DEC_M: SX_PK_EACH(s.m = (s.m - one) & ff; pk_set_nz(s, s.m))
  case ASL_M: SX_PK_EACH(s.m = pk_asl(s, s.m))
  case LSR_M: SX_PK_EACH(s.m = pk_lsr(s, s.m))
  case ROL_M: SX_PK_EACH(s.m = pk_rol(s, s.m))
  case ROR_M: SX_PK_EACH(s.m = pk_ror(s, s.m))
  case BIT_M:
    SX_PK_EACH(s.p = (s.p & pk_splat(0x3D)) | (s.m & pk_splat(0xC0)) |
                     ((((s.a & s.m) - one) >> 7) & pk_splat(2)))
  default: return false;
  }
  return true;
}
#undef SX_PK_EACH

}  // namespace sx
