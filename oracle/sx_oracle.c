/*
 * sx_oracle.c -- CPU oracle for the 6502 search core (TEST INFRASTRUCTURE).
 *
 * A plain-C restatement of the Go algorithm of oisee/6502-optimizer, written
 * from the Go sources read as text.  Only tests/, the smoke check and the
 * bench's CPU-baseline leg load it, always as the checker.  The product
 * library (6502-optimizer_amd/csrc) never links or calls this file.
 *
 * Citations are path:line in the reference tree.
 */
#include "sx_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ ISA -- */

/* pkg/inst/instruction.go:13-21 */
int sxo_has_imm(int op) { return op >= SXO_LDA_IMM && op <= SXO_CPY_IMM; }

/* pkg/inst/catalog.go:105-201 (mnemonics, encodings, cycles). */
static const char *const k_mnem[SXO_OPCOUNT] = {
  "TAX", "TAY", "TXA", "TYA", "TSX", "TXS",
  "INX", "INY", "DEX", "DEY",
  "ASL A", "LSR A", "ROL A", "ROR A",
  "CLC", "SEC", "CLD", "SED", "CLI", "SEI", "CLV",
  "PHA", "PLA", "PHP", "PLP", "NOP",
  "LDA #n", "LDX #n", "LDY #n", "ADC #n", "SBC #n",
  "AND #n", "ORA #n", "EOR #n", "CMP #n", "CPX #n", "CPY #n",
  "LDA M", "LDX M", "LDY M", "STA M", "STX M", "STY M",
  "ADC M", "SBC M", "AND M", "ORA M", "EOR M",
  "CMP M", "CPX M", "CPY M", "INC M", "DEC M",
  "ASL M", "LSR M", "ROL M", "ROR M", "BIT M",
};

const char *sxo_mnemonic(int op) {
  return (op >= 0 && op < SXO_OPCOUNT) ? k_mnem[op] : NULL;
}

/* Go ByteSize (catalog.go:50-56): one encoding byte, plus one for an
 * immediate, so the 21 M ops count 1 byte.  The CUDA tool counts every
 * non-implied op as 2 bytes (cuda/6502_search_v2.cu:437-440). */
int sxo_byte_size(int op, int model) {
  if (model == 1) return op <= SXO_NOP ? 1 : 2;
  return sxo_has_imm(op) ? 2 : 1;
}

/* catalog.go:108-200 cycle column. */
int sxo_cycles(int op) {
  switch (op) {
  case SXO_PHA: case SXO_PHP: return 3;
  case SXO_PLA: case SXO_PLP: return 4;
  case SXO_INC_M: case SXO_DEC_M: case SXO_ASL_M: case SXO_LSR_M:
  case SXO_ROL_M: case SXO_ROR_M: return 5;
  default: break;
  }
  if (op >= SXO_LDA_M && op <= SXO_BIT_M) return 3;
  return 2;
}

/* catalog.go:59-85: the 'n' placeholder becomes $XX in upper-case hex. */
int sxo_disasm(uint32_t w, char *buf, int cap) {
  int op = SXO_OP(w);
  if (op >= SXO_OPCOUNT) return snprintf(buf, cap, "OP(%d)", op);
  const char *m = k_mnem[op];
  if (!sxo_has_imm(op)) return snprintf(buf, cap, "%s", m);
  static const char hx[] = "0123456789ABCDEF";
  char tmp[32];
  int o = 0;
  for (int i = 0; m[i]; i++) {
    if (m[i] == 'n' && (i == 0 || m[i - 1] == '#' || m[i - 1] == ' ')) {
      uint8_t v = (uint8_t)SXO_IMM(w);
      tmp[o++] = '$';
      tmp[o++] = hx[v >> 4];
      tmp[o++] = hx[v & 15];
    } else {
      tmp[o++] = m[i];
    }
  }
  tmp[o] = 0;
  return snprintf(buf, cap, "%s", tmp);
}

/* Go order: enumerator.go:14-38 (non-immediate ops in opcode order, then each
 * immediate op with imm 0..255).  CUDA order: cuda/6502_search_v2.cu:741-752
 * (opcode order with immediates inline). */
int sxo_instruction_table(uint32_t *out, int order) {
  int n = 0;
  if (order == 0) {
    for (int op = 0; op < SXO_OPCOUNT; op++)
      if (!sxo_has_imm(op)) out[n++] = SXO_WORD(op, 0);
    for (int op = 0; op < SXO_OPCOUNT; op++)
      if (sxo_has_imm(op))
        for (int imm = 0; imm < 256; imm++) out[n++] = SXO_WORD(op, imm);
  } else {
    for (int op = 0; op < SXO_OPCOUNT; op++) {
      if (sxo_has_imm(op)) {
        for (int imm = 0; imm < 256; imm++) out[n++] = SXO_WORD(op, imm);
      } else {
        out[n++] = SXO_WORD(op, 0);
      }
    }
  }
  return n;
}

/* ------------------------------------------------------------- Executor -- */

#define FC 0x01
#define FZ 0x02
#define FI 0x04
#define FD 0x08
#define FB 0x10
#define FU 0x20
#define FV 0x40
#define FN 0x80
#define PMASK 0xCF

/* pkg/cpu/flags.go:17-30 */
static inline uint8_t nz(uint8_t v) { return (uint8_t)((v == 0 ? FZ : 0) | (v & 0x80 ? FN : 0)); }

/* exec.go:194-197 */
static inline void set_nz(uint8_t *s, uint8_t v) {
  s[SXO_P] = (uint8_t)((s[SXO_P] & ~(FN | FZ)) | nz(v));
}

/* exec.go:202-220 */
static inline void do_adc(uint8_t *s, uint8_t m) {
  unsigned a = s[SXO_A], c = s[SXO_P] & FC;
  unsigned r = a + m + c;
  uint8_t v = (uint8_t)((a ^ (uint8_t)r) & (m ^ (uint8_t)r) & 0x80);
  s[SXO_A] = (uint8_t)r;
  s[SXO_P] &= (uint8_t)~(FC | FZ | FN | FV);
  s[SXO_P] |= nz(s[SXO_A]);
  if (r > 0xFF) s[SXO_P] |= FC;
  if (v) s[SXO_P] |= FV;
}

/* exec.go:224-242: uint16 arithmetic, so a borrow wraps above 0xFF. */
static inline void do_sbc(uint8_t *s, uint8_t m) {
  uint16_t a = s[SXO_A], c = s[SXO_P] & FC;
  uint16_t r = (uint16_t)(a - m - (1 - c));
  uint8_t v = (uint8_t)((s[SXO_A] ^ (uint8_t)r) & (s[SXO_A] ^ m) & 0x80);
  s[SXO_A] = (uint8_t)r;
  s[SXO_P] &= (uint8_t)~(FC | FZ | FN | FV);
  s[SXO_P] |= nz(s[SXO_A]);
  if (r <= 0xFF) s[SXO_P] |= FC;
  if (v) s[SXO_P] |= FV;
}

/* exec.go:246-253 */
static inline void do_cmp(uint8_t *s, uint8_t reg, uint8_t m) {
  uint16_t r = (uint16_t)((uint16_t)reg - (uint16_t)m);
  s[SXO_P] &= (uint8_t)~(FC | FZ | FN);
  s[SXO_P] |= nz((uint8_t)r);
  if (reg >= m) s[SXO_P] |= FC;
}

/* exec.go:256-294 */
static inline void do_asl(uint8_t *s, uint8_t *v) {
  s[SXO_P] &= (uint8_t)~FC;
  if (*v & 0x80) s[SXO_P] |= FC;
  *v = (uint8_t)(*v << 1);
  set_nz(s, *v);
}
static inline void do_lsr(uint8_t *s, uint8_t *v) {
  s[SXO_P] &= (uint8_t)~FC;
  if (*v & 0x01) s[SXO_P] |= FC;
  *v = (uint8_t)(*v >> 1);
  set_nz(s, *v);
}
static inline void do_rol(uint8_t *s, uint8_t *v) {
  uint8_t oc = s[SXO_P] & FC;
  s[SXO_P] &= (uint8_t)~FC;
  if (*v & 0x80) s[SXO_P] |= FC;
  *v = (uint8_t)((*v << 1) | oc);
  set_nz(s, *v);
}
static inline void do_ror(uint8_t *s, uint8_t *v) {
  uint8_t oc = s[SXO_P] & FC;
  s[SXO_P] &= (uint8_t)~FC;
  if (*v & 0x01) s[SXO_P] |= FC;
  *v = (uint8_t)((*v >> 1) | (oc << 7));
  set_nz(s, *v);
}

/* exec.go:298-310 */
static inline void do_bit(uint8_t *s, uint8_t m) {
  s[SXO_P] &= (uint8_t)~(FN | FV | FZ);
  if (m & 0x80) s[SXO_P] |= FN;
  if (m & 0x40) s[SXO_P] |= FV;
  if ((s[SXO_A] & m) == 0) s[SXO_P] |= FZ;
}

/* exec.go:314-327: two-deep virtual stack. */
static inline void do_push(uint8_t *s, uint8_t v) {
  s[SXO_S1] = s[SXO_S0];
  s[SXO_S0] = v;
  s[SXO_S]--;
}
static inline uint8_t do_pull(uint8_t *s) {
  uint8_t v = s[SXO_S0];
  s[SXO_S0] = s[SXO_S1];
  s[SXO_S1] = 0;
  s[SXO_S]++;
  return v;
}

/* pkg/cpu/exec.go:8-189 */
static inline __attribute__((always_inline)) int exec_one(uint8_t *s, int op, uint16_t imm16) {
  uint8_t imm = (uint8_t)imm16;
  switch (op) {
  case SXO_TAX: s[SXO_X] = s[SXO_A]; set_nz(s, s[SXO_X]); break;
  case SXO_TAY: s[SXO_Y] = s[SXO_A]; set_nz(s, s[SXO_Y]); break;
  case SXO_TXA: s[SXO_A] = s[SXO_X]; set_nz(s, s[SXO_A]); break;
  case SXO_TYA: s[SXO_A] = s[SXO_Y]; set_nz(s, s[SXO_A]); break;
  case SXO_TSX: s[SXO_X] = s[SXO_S]; set_nz(s, s[SXO_X]); break;
  case SXO_TXS: s[SXO_S] = s[SXO_X]; break;
  case SXO_INX: s[SXO_X]++; set_nz(s, s[SXO_X]); break;
  case SXO_INY: s[SXO_Y]++; set_nz(s, s[SXO_Y]); break;
  case SXO_DEX: s[SXO_X]--; set_nz(s, s[SXO_X]); break;
  case SXO_DEY: s[SXO_Y]--; set_nz(s, s[SXO_Y]); break;
  case SXO_ASL_A: do_asl(s, &s[SXO_A]); break;
  case SXO_LSR_A: do_lsr(s, &s[SXO_A]); break;
  case SXO_ROL_A: do_rol(s, &s[SXO_A]); break;
  case SXO_ROR_A: do_ror(s, &s[SXO_A]); break;
  case SXO_CLC: s[SXO_P] &= (uint8_t)~FC; break;
  case SXO_SEC: s[SXO_P] |= FC; break;
  case SXO_CLD: s[SXO_P] &= (uint8_t)~FD; break;
  case SXO_SED: s[SXO_P] |= FD; break;
  case SXO_CLI: s[SXO_P] &= (uint8_t)~FI; break;
  case SXO_SEI: s[SXO_P] |= FI; break;
  case SXO_CLV: s[SXO_P] &= (uint8_t)~FV; break;
  case SXO_PHA: do_push(s, s[SXO_A]); break;
  case SXO_PLA: s[SXO_A] = do_pull(s); set_nz(s, s[SXO_A]); break;
  case SXO_PHP: do_push(s, (uint8_t)(s[SXO_P] | FB | FU)); break;
  case SXO_PLP: s[SXO_P] = do_pull(s); break;
  case SXO_NOP: break;
  case SXO_LDA_IMM: s[SXO_A] = imm; set_nz(s, s[SXO_A]); break;
  case SXO_LDX_IMM: s[SXO_X] = imm; set_nz(s, s[SXO_X]); break;
  case SXO_LDY_IMM: s[SXO_Y] = imm; set_nz(s, s[SXO_Y]); break;
  case SXO_ADC_IMM: do_adc(s, imm); break;
  case SXO_SBC_IMM: do_sbc(s, imm); break;
  case SXO_AND_IMM: s[SXO_A] &= imm; set_nz(s, s[SXO_A]); break;
  case SXO_ORA_IMM: s[SXO_A] |= imm; set_nz(s, s[SXO_A]); break;
  case SXO_EOR_IMM: s[SXO_A] ^= imm; set_nz(s, s[SXO_A]); break;
  case SXO_CMP_IMM: do_cmp(s, s[SXO_A], imm); break;
  case SXO_CPX_IMM: do_cmp(s, s[SXO_X], imm); break;
  case SXO_CPY_IMM: do_cmp(s, s[SXO_Y], imm); break;
  case SXO_LDA_M: s[SXO_A] = s[SXO_M]; set_nz(s, s[SXO_A]); break;
  case SXO_LDX_M: s[SXO_X] = s[SXO_M]; set_nz(s, s[SXO_X]); break;
  case SXO_LDY_M: s[SXO_Y] = s[SXO_M]; set_nz(s, s[SXO_Y]); break;
  case SXO_STA_M: s[SXO_M] = s[SXO_A]; break;
  case SXO_STX_M: s[SXO_M] = s[SXO_X]; break;
  case SXO_STY_M: s[SXO_M] = s[SXO_Y]; break;
  case SXO_ADC_M: do_adc(s, s[SXO_M]); break;
  case SXO_SBC_M: do_sbc(s, s[SXO_M]); break;
  case SXO_AND_M: s[SXO_A] &= s[SXO_M]; set_nz(s, s[SXO_A]); break;
  case SXO_ORA_M: s[SXO_A] |= s[SXO_M]; set_nz(s, s[SXO_A]); break;
  case SXO_EOR_M: s[SXO_A] ^= s[SXO_M]; set_nz(s, s[SXO_A]); break;
  case SXO_CMP_M: do_cmp(s, s[SXO_A], s[SXO_M]); break;
  case SXO_CPX_M: do_cmp(s, s[SXO_X], s[SXO_M]); break;
  case SXO_CPY_M: do_cmp(s, s[SXO_Y], s[SXO_M]); break;
  case SXO_INC_M: s[SXO_M]++; set_nz(s, s[SXO_M]); break;
  case SXO_DEC_M: s[SXO_M]--; set_nz(s, s[SXO_M]); break;
  case SXO_ASL_M: do_asl(s, &s[SXO_M]); break;
  case SXO_LSR_M: do_lsr(s, &s[SXO_M]); break;
  case SXO_ROL_M: do_rol(s, &s[SXO_M]); break;
  case SXO_ROR_M: do_ror(s, &s[SXO_M]); break;
  case SXO_BIT_M: do_bit(s, s[SXO_M]); break;
  default: return -1; /* Go panics here (exec.go:185-186) */
  }
  return sxo_cycles(op);
}

int sxo_exec(uint8_t *s, int op, uint16_t imm16) { return exec_one(s, op, imm16); }

void sxo_exec_many(uint8_t *states, const uint32_t *words, int64_t n) {
  for (int64_t i = 0; i < n; i++) exec_one(states + 8 * i, SXO_OP(words[i]), SXO_IMM(words[i]));
}

/* verifier.go:70-76 */
void sxo_exec_seq(uint8_t *s, const uint32_t *seq, int n) {
  for (int i = 0; i < n; i++) exec_one(s, SXO_OP(seq[i]), SXO_IMM(seq[i]));
}

/* --------------------------------------------------------------- Vectors -- */

/* verifier.go:18-27, A X Y P S M S0 S1. */
static const uint8_t k_qc[8][8] = {
  {0x00, 0x00, 0x00, 0x00, 0xFD, 0x42, 0x00, 0x00},
  {0xFF, 0xFF, 0xFF, 0xFF, 0xFF, 0xBD, 0xFF, 0xFF},
  {0x01, 0x02, 0x03, 0x00, 0xFD, 0x13, 0x10, 0x20},
  {0x80, 0x40, 0x20, 0x01, 0xFC, 0x7E, 0x30, 0x40},
  {0x55, 0xAA, 0x55, 0x00, 0xFE, 0xA5, 0x55, 0xAA},
  {0xAA, 0x55, 0xAA, 0x01, 0xFD, 0x5A, 0xAA, 0x55},
  {0x0F, 0xF0, 0x0F, 0x00, 0xFB, 0xE1, 0x0F, 0xF0},
  {0x7F, 0x80, 0x7F, 0x01, 0xFD, 0x33, 0x80, 0x7F},
};

/* verifier.go:30-67 (the first 8 equal the QuickCheck vectors). */
static const uint8_t k_mid[32][8] = {
  {0x00, 0x00, 0x00, 0x00, 0xFD, 0x42, 0x00, 0x00},
  {0xFF, 0xFF, 0xFF, 0xFF, 0xFF, 0xBD, 0xFF, 0xFF},
  {0x01, 0x02, 0x03, 0x00, 0xFD, 0x13, 0x10, 0x20},
  {0x80, 0x40, 0x20, 0x01, 0xFC, 0x7E, 0x30, 0x40},
  {0x55, 0xAA, 0x55, 0x00, 0xFE, 0xA5, 0x55, 0xAA},
  {0xAA, 0x55, 0xAA, 0x01, 0xFD, 0x5A, 0xAA, 0x55},
  {0x0F, 0xF0, 0x0F, 0x00, 0xFB, 0xE1, 0x0F, 0xF0},
  {0x7F, 0x80, 0x7F, 0x01, 0xFD, 0x33, 0x80, 0x7F},
  {0x02, 0x01, 0x04, 0x00, 0xFD, 0x91, 0x11, 0x22},
  {0x04, 0x08, 0x10, 0x01, 0xFD, 0x48, 0x33, 0x44},
  {0x08, 0x10, 0x20, 0x00, 0xFD, 0x24, 0x55, 0x66},
  {0x10, 0x20, 0x40, 0x01, 0xFD, 0xD7, 0x77, 0x88},
  {0x20, 0x40, 0x80, 0x00, 0xFD, 0x6B, 0x99, 0xBB},
  {0x40, 0x80, 0x01, 0x01, 0xFD, 0xB2, 0xCC, 0xDD},
  {0x03, 0x05, 0x09, 0x00, 0xFD, 0x0F, 0xEE, 0x11},
  {0x05, 0x0A, 0x14, 0x01, 0xFD, 0xF0, 0x22, 0x33},
  {0xFE, 0x7F, 0x80, 0x00, 0xFD, 0x01, 0x44, 0x55},
  {0x81, 0xFE, 0x01, 0x01, 0xFD, 0xFE, 0x66, 0x77},
  {0x7E, 0x81, 0x42, 0x00, 0xFD, 0x80, 0x88, 0x99},
  {0xBF, 0x40, 0xBF, 0x01, 0xFD, 0x40, 0xAA, 0xBB},
  {0xC0, 0x3F, 0xC0, 0x00, 0xFD, 0x1F, 0xCC, 0xDD},
  {0xE0, 0x1F, 0xE0, 0x01, 0xFD, 0xEF, 0xEE, 0x00},
  {0xF0, 0x0F, 0xF0, 0x00, 0xFD, 0x55, 0x11, 0x22},
  {0x0E, 0xE0, 0x0E, 0x01, 0xFD, 0xAA, 0x33, 0x44},
  {0x11, 0x22, 0x33, 0x00, 0xFE, 0xC3, 0x00, 0x00},
  {0x22, 0x33, 0x44, 0x01, 0xFF, 0x3C, 0x55, 0x66},
  {0x44, 0x55, 0x66, 0x00, 0xFC, 0x87, 0x77, 0x88},
  {0x88, 0x99, 0xAA, 0x01, 0xFB, 0x69, 0x99, 0xAA},
  {0x33, 0x44, 0x55, 0x00, 0xFD, 0xAE, 0xBB, 0xCC},
  {0xCC, 0xDD, 0xEE, 0x01, 0xFD, 0x51, 0xDD, 0xEE},
  {0x66, 0x77, 0x88, 0x00, 0xFD, 0xDA, 0x12, 0x34},
  {0x99, 0xAA, 0xBB, 0x01, 0xFD, 0x25, 0x56, 0x78},
};

const uint8_t *sxo_qc_vectors(void) { return &k_qc[0][0]; }
const uint8_t *sxo_mid_vectors(void) { return &k_mid[0][0]; }

/* -------------------------------------------------------------- Verifier -- */

/* verifier.go:418-424 (dead = 0 reduces to State.Equal, state.go:23-28). */
int sxo_states_equal_masked(const uint8_t *a, const uint8_t *b, uint8_t dead) {
  uint8_t pm = (uint8_t)(PMASK & ~dead);
  return a[SXO_A] == b[SXO_A] && a[SXO_X] == b[SXO_X] && a[SXO_Y] == b[SXO_Y] &&
         (a[SXO_P] & pm) == (b[SXO_P] & pm) && a[SXO_S] == b[SXO_S] &&
         a[SXO_M] == b[SXO_M] && a[SXO_S0] == b[SXO_S0] && a[SXO_S1] == b[SXO_S1];
}

static int check_vectors(const uint8_t (*vec)[8], int nv, const uint32_t *t, int tn,
                         const uint32_t *c, int cn, uint8_t dead) {
  for (int i = 0; i < nv; i++) {
    uint8_t ts[8], cs[8];
    memcpy(ts, vec[i], 8);
    memcpy(cs, vec[i], 8);
    sxo_exec_seq(ts, t, tn);
    sxo_exec_seq(cs, c, cn);
    if (!sxo_states_equal_masked(ts, cs, dead)) return 0;
  }
  return 1;
}

/* verifier.go:84-93 and :427-439 */
int sxo_quick_check(const uint32_t *t, int tn, const uint32_t *c, int cn, uint8_t dead) {
  return check_vectors(k_qc, 8, t, tn, c, cn, dead);
}

/* verifier.go:96-105 and :442-454 */
int sxo_mid_check(const uint32_t *t, int tn, const uint32_t *c, int cn, uint8_t dead) {
  return check_vectors(k_mid, 32, t, tn, c, cn, dead);
}

/* verifier.go:114-129 */
void sxo_fingerprint(const uint32_t *seq, int n, uint8_t fp[64]) {
  for (int i = 0; i < 8; i++) {
    uint8_t s[8];
    memcpy(s, k_qc[i], 8);
    sxo_exec_seq(s, seq, n);
    memcpy(fp + 8 * i, s, 8);
    fp[8 * i + SXO_P] &= PMASK;
  }
}

/* verifier.go:591-604 */
uint8_t sxo_flag_diff(const uint32_t *t, int tn, const uint32_t *c, int cn) {
  uint8_t diff = 0;
  for (int i = 0; i < 8; i++) {
    uint8_t ts[8], cs[8];
    memcpy(ts, k_qc[i], 8);
    memcpy(cs, k_qc[i], 8);
    sxo_exec_seq(ts, t, tn);
    sxo_exec_seq(cs, c, cn);
    for (int r = 0; r < 8; r++)
      if (r != SXO_P && ts[r] != cs[r]) return 0;
    diff |= (uint8_t)((ts[SXO_P] ^ cs[SXO_P]) & PMASK);
  }
  return diff;
}

/* verifier.go:154-259 */
int sxo_op_reads(int op) {
  switch (op) {
  case SXO_TAX: case SXO_TAY: return SXO_RA;
  case SXO_TXA: return SXO_RX;
  case SXO_TYA: return SXO_RY;
  case SXO_TSX: return SXO_RS;
  case SXO_TXS: return SXO_RX;
  case SXO_INX: case SXO_DEX: return SXO_RX;
  case SXO_INY: case SXO_DEY: return SXO_RY;
  case SXO_ASL_A: case SXO_LSR_A: return SXO_RA;
  case SXO_ROL_A: case SXO_ROR_A: return SXO_RA | SXO_RP;
  case SXO_CLC: case SXO_SEC: case SXO_CLD: case SXO_SED:
  case SXO_CLI: case SXO_SEI: case SXO_CLV: return 0;
  case SXO_PHA: return SXO_RA | SXO_RS0 | SXO_RS1;
  case SXO_PLA: return SXO_RS0;
  case SXO_PHP: return SXO_RP | SXO_RS0 | SXO_RS1;
  case SXO_PLP: return SXO_RS0;
  case SXO_NOP: return 0;
  case SXO_LDA_IMM: case SXO_LDX_IMM: case SXO_LDY_IMM: return 0;
  case SXO_ADC_IMM: case SXO_SBC_IMM: return SXO_RA | SXO_RP;
  case SXO_AND_IMM: case SXO_ORA_IMM: case SXO_EOR_IMM: return SXO_RA;
  case SXO_CMP_IMM: return SXO_RA;
  case SXO_CPX_IMM: return SXO_RX;
  case SXO_CPY_IMM: return SXO_RY;
  case SXO_LDA_M: case SXO_LDX_M: case SXO_LDY_M: return SXO_RM;
  case SXO_STA_M: return SXO_RA;
  case SXO_STX_M: return SXO_RX;
  case SXO_STY_M: return SXO_RY;
  case SXO_ADC_M: case SXO_SBC_M: return SXO_RA | SXO_RM | SXO_RP;
  case SXO_AND_M: case SXO_ORA_M: case SXO_EOR_M: return SXO_RA | SXO_RM;
  case SXO_CMP_M: return SXO_RA | SXO_RM;
  case SXO_CPX_M: return SXO_RX | SXO_RM;
  case SXO_CPY_M: return SXO_RY | SXO_RM;
  case SXO_INC_M: case SXO_DEC_M: return SXO_RM;
  case SXO_ASL_M: case SXO_LSR_M: return SXO_RM;
  case SXO_ROL_M: case SXO_ROR_M: return SXO_RM | SXO_RP;
  case SXO_BIT_M: return SXO_RA | SXO_RM;
  }
  return 0;
}

/* pruner.go:76-163 */
int sxo_op_writes(int op) {
  switch (op) {
  case SXO_TAX: return SXO_RX | SXO_RP;
  case SXO_TAY: return SXO_RY | SXO_RP;
  case SXO_TXA: case SXO_TYA: return SXO_RA | SXO_RP;
  case SXO_TSX: return SXO_RX | SXO_RP;
  case SXO_TXS: return SXO_RS;
  case SXO_INX: case SXO_DEX: return SXO_RX | SXO_RP;
  case SXO_INY: case SXO_DEY: return SXO_RY | SXO_RP;
  case SXO_ASL_A: case SXO_LSR_A: case SXO_ROL_A: case SXO_ROR_A: return SXO_RA | SXO_RP;
  case SXO_CLC: case SXO_SEC: case SXO_CLD: case SXO_SED:
  case SXO_CLI: case SXO_SEI: case SXO_CLV: return SXO_RP;
  case SXO_PHA: case SXO_PHP: return SXO_RS | SXO_RS0 | SXO_RS1;
  case SXO_PLA: return SXO_RA | SXO_RP | SXO_RS | SXO_RS0;
  case SXO_PLP: return SXO_RP | SXO_RS | SXO_RS0;
  case SXO_NOP: return 0;
  case SXO_LDA_IMM: return SXO_RA | SXO_RP;
  case SXO_LDX_IMM: return SXO_RX | SXO_RP;
  case SXO_LDY_IMM: return SXO_RY | SXO_RP;
  case SXO_ADC_IMM: case SXO_SBC_IMM: case SXO_AND_IMM: case SXO_ORA_IMM:
  case SXO_EOR_IMM: return SXO_RA | SXO_RP;
  case SXO_CMP_IMM: case SXO_CPX_IMM: case SXO_CPY_IMM: return SXO_RP;
  case SXO_LDA_M: return SXO_RA | SXO_RP;
  case SXO_LDX_M: return SXO_RX | SXO_RP;
  case SXO_LDY_M: return SXO_RY | SXO_RP;
  case SXO_STA_M: case SXO_STX_M: case SXO_STY_M: return SXO_RM;
  case SXO_ADC_M: case SXO_SBC_M: case SXO_AND_M: case SXO_ORA_M:
  case SXO_EOR_M: return SXO_RA | SXO_RP;
  case SXO_CMP_M: case SXO_CPX_M: case SXO_CPY_M: return SXO_RP;
  case SXO_INC_M: case SXO_DEC_M: return SXO_RM | SXO_RP;
  case SXO_ASL_M: case SXO_LSR_M: case SXO_ROL_M: case SXO_ROR_M: return SXO_RM | SXO_RP;
  case SXO_BIT_M: return SXO_RP;
  }
  return 0;
}

/* verifier.go:145-151 */
int sxo_regs_read(const uint32_t *seq, int n) {
  int m = 0;
  for (int i = 0; i < n; i++) m |= sxo_op_reads(SXO_OP(seq[i]));
  return m;
}

/* verifier.go:292-310: extra registers in the order X, Y, M, S, S0, S1,
 * as setReg offsets 1..6 (verifier.go:322-337). */
int sxo_extra_regs(int reads, int out[6]) {
  int k = 0;
  if (reads & SXO_RX) out[k++] = 1;
  if (reads & SXO_RY) out[k++] = 2;
  if (reads & SXO_RM) out[k++] = 3;
  if (reads & SXO_RS) out[k++] = 4;
  if (reads & SXO_RS0) out[k++] = 5;
  if (reads & SXO_RS1) out[k++] = 6;
  return k;
}

/* setReg offset -> state byte index */
static const int k_off_to_byte[7] = {SXO_A, SXO_X, SXO_Y, SXO_M, SXO_S, SXO_S0, SXO_S1};

/* verifier.go:378-383 */
static const uint8_t k_rep[32] = {
  0x00, 0x01, 0x02, 0x0F, 0x10, 0x1F, 0x20, 0x3F,
  0x40, 0x55, 0x7E, 0x7F, 0x80, 0x81, 0xAA, 0xBF,
  0xC0, 0xD5, 0xE0, 0xEF, 0xF0, 0xF7, 0xFE, 0xFF,
  0x03, 0x07, 0x11, 0x33, 0x77, 0xBB, 0xDD, 0xEE,
};

typedef union st8 {
  uint8_t b[8];
  uint64_t q;
} st8;

/* Masked compare of two states: every byte, P under PMask &^ dead
 * (verifier.go:418-425). */
static inline int st_equal(const st8 *a, const st8 *b, uint8_t dead) {
  st8 msk;
  msk.q = ~(uint64_t)0;
  msk.b[SXO_P] = (uint8_t)(PMASK & ~dead);
  return ((a->q ^ b->q) & msk.q) == 0;
}

/* Run both sequences on one state; 1 when the outputs agree. */
static inline int pair_agrees(const st8 *in, const uint32_t *t, int tn, const uint32_t *c, int cn,
                              uint8_t dead) {
  st8 ts = *in, cs = *in;
  for (int i = 0; i < tn; i++) exec_one(ts.b, SXO_OP(t[i]), SXO_IMM(t[i]));
  for (int i = 0; i < cn; i++) exec_one(cs.b, SXO_OP(c[i]), SXO_IMM(c[i]));
  return st_equal(&ts, &cs, dead);
}

/* The domain of verifier.go:262-415: A = 0..255, P = carry 0/1 (other flags
 * 0), S = 0xFD unless read, X, Y, M, S0, S1 = 0 unless read; the registers
 * either side reads (opReads) are swept on top.  Returns the state bytes it
 * varies as a bit set (bit r = state byte r; P stands for the carry). */
static int domain_bytes(const uint32_t *t, int tn, const uint32_t *c, int cn) {
  int reads = sxo_regs_read(t, tn) | sxo_regs_read(c, cn);
  int v = (1 << SXO_A) | (1 << SXO_P);
  if (reads & SXO_RX) v |= 1 << SXO_X;
  if (reads & SXO_RY) v |= 1 << SXO_Y;
  if (reads & SXO_RM) v |= 1 << SXO_M;
  if (reads & SXO_RS) v |= 1 << SXO_S;
  if (reads & SXO_RS0) v |= 1 << SXO_S0;
  if (reads & SXO_RS1) v |= 1 << SXO_S1;
  return v;
}

/* Sweep the bytes of `set` (bit r = state byte r; P = carry 0/1) over their
 * domain values with every other byte at its domain default; 1 when T and C
 * agree everywhere, else 0 with the first disagreeing state in cex. */
static int sweep_set(const uint32_t *t, int tn, const uint32_t *c, int cn, uint8_t dead, int set,
                     uint8_t cex[8]) {
  int regs[8], k = 0;
  for (int r = 0; r < 8; r++)
    if (set & (1 << r)) regs[k++] = r;
  int odo[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  st8 s;
  s.q = 0;
  s.b[SXO_S] = 0xFD;
  for (;;) {
    for (int i = 0; i < k; i++) s.b[regs[i]] = (uint8_t)odo[i];
    if (!pair_agrees(&s, t, tn, c, cn, dead)) {
      if (cex) memcpy(cex, s.b, 8);
      return 0;
    }
    int i = 0;
    while (i < k && ++odo[i] == (regs[i] == SXO_P ? 2 : 256)) odo[i++] = 0;
    if (i == k) return 1;
  }
}

/* exec_one over a block of states with one opcode dispatch per block: each
 * case calls the (inlined) executor with a constant opcode, so its switch
 * folds away and the loop runs that opcode's code only.  Same semantics as
 * exec_one by construction. */
#define BLOCK_CASE(OPC) \
  case OPC: for (int i = 0; i < n; i++) exec_one(s[i].b, OPC, imm); break;
static void exec_block(st8 *s, int n, int op, uint16_t imm) {
  switch (op) {
  BLOCK_CASE(SXO_TAX) BLOCK_CASE(SXO_TAY) BLOCK_CASE(SXO_TXA) BLOCK_CASE(SXO_TYA) BLOCK_CASE(SXO_TSX)
  BLOCK_CASE(SXO_TXS) BLOCK_CASE(SXO_INX) BLOCK_CASE(SXO_INY) BLOCK_CASE(SXO_DEX) BLOCK_CASE(SXO_DEY)
  BLOCK_CASE(SXO_ASL_A) BLOCK_CASE(SXO_LSR_A) BLOCK_CASE(SXO_ROL_A) BLOCK_CASE(SXO_ROR_A)
  BLOCK_CASE(SXO_CLC) BLOCK_CASE(SXO_SEC) BLOCK_CASE(SXO_CLD) BLOCK_CASE(SXO_SED) BLOCK_CASE(SXO_CLI)
  BLOCK_CASE(SXO_SEI) BLOCK_CASE(SXO_CLV) BLOCK_CASE(SXO_PHA) BLOCK_CASE(SXO_PLA) BLOCK_CASE(SXO_PHP)
  BLOCK_CASE(SXO_PLP) BLOCK_CASE(SXO_NOP) BLOCK_CASE(SXO_LDA_IMM) BLOCK_CASE(SXO_LDX_IMM)
  BLOCK_CASE(SXO_LDY_IMM) BLOCK_CASE(SXO_ADC_IMM) BLOCK_CASE(SXO_SBC_IMM) BLOCK_CASE(SXO_AND_IMM)
  BLOCK_CASE(SXO_ORA_IMM) BLOCK_CASE(SXO_EOR_IMM) BLOCK_CASE(SXO_CMP_IMM) BLOCK_CASE(SXO_CPX_IMM)
  BLOCK_CASE(SXO_CPY_IMM) BLOCK_CASE(SXO_LDA_M) BLOCK_CASE(SXO_LDX_M) BLOCK_CASE(SXO_LDY_M)
  BLOCK_CASE(SXO_STA_M) BLOCK_CASE(SXO_STX_M) BLOCK_CASE(SXO_STY_M) BLOCK_CASE(SXO_ADC_M)
  BLOCK_CASE(SXO_SBC_M) BLOCK_CASE(SXO_AND_M) BLOCK_CASE(SXO_ORA_M) BLOCK_CASE(SXO_EOR_M)
  BLOCK_CASE(SXO_CMP_M) BLOCK_CASE(SXO_CPX_M) BLOCK_CASE(SXO_CPY_M) BLOCK_CASE(SXO_INC_M)
  BLOCK_CASE(SXO_DEC_M) BLOCK_CASE(SXO_ASL_M) BLOCK_CASE(SXO_LSR_M) BLOCK_CASE(SXO_ROL_M)
  BLOCK_CASE(SXO_ROR_M) BLOCK_CASE(SXO_BIT_M)
  default: break;
  }
}
#undef BLOCK_CASE

void sxo_exec_block(uint8_t *states, int n, uint32_t w) {
  exec_block((st8 *)states, n, SXO_OP(w), SXO_IMM(w));
}

/* Complete brute force over the whole domain (mode 1) with block execution:
 * the 512 states of A x carry are one block, the extra registers an odometer
 * split over nthreads threads.  1 = pass; else 0 and, in cex, a state where
 * the outputs differ (the first one in the odometer order of its thread). */
typedef struct brute_ctx {
  const uint32_t *t, *c;
  int tn, cn;
  uint8_t dead;
  int regs[6], k;
  uint64_t nouter, next;
  int failed;
  uint8_t cex[8];
  pthread_mutex_t mu;
} brute_ctx;

static void *brute_worker(void *arg) {
  brute_ctx *w = (brute_ctx *)arg;
  st8 ts[512], cs[512], base[512];
  for (int i = 0; i < 512; i++) {
    base[i].q = 0;
    base[i].b[SXO_A] = (uint8_t)(i & 255);
    base[i].b[SXO_P] = (uint8_t)(i >> 8);
    base[i].b[SXO_S] = 0xFD;
  }
  st8 msk;
  msk.q = ~(uint64_t)0;
  msk.b[SXO_P] = (uint8_t)(PMASK & ~w->dead);
  const uint64_t chunk = 256;
  for (;;) {
    if (__atomic_load_n(&w->failed, __ATOMIC_RELAXED)) break;
    uint64_t b = __atomic_fetch_add(&w->next, chunk, __ATOMIC_RELAXED);
    if (b >= w->nouter) break;
    uint64_t e = b + chunk < w->nouter ? b + chunk : w->nouter;
    for (uint64_t o = b; o < e; o++) {
      for (int i = 0; i < w->k; i++) {
        const uint8_t v = (uint8_t)(o >> (8 * i));
        for (int j = 0; j < 512; j++) base[j].b[w->regs[i]] = v;
      }
      memcpy(ts, base, sizeof(ts));
      memcpy(cs, base, sizeof(cs));
      for (int q = 0; q < w->tn; q++) exec_block(ts, 512, SXO_OP(w->t[q]), SXO_IMM(w->t[q]));
      for (int q = 0; q < w->cn; q++) exec_block(cs, 512, SXO_OP(w->c[q]), SXO_IMM(w->c[q]));
      uint64_t bad = 0;
      for (int j = 0; j < 512; j++) bad |= (ts[j].q ^ cs[j].q) & msk.q;
      if (bad) {
        for (int j = 0; j < 512; j++)
          if ((ts[j].q ^ cs[j].q) & msk.q) {
            pthread_mutex_lock(&w->mu);
            if (!w->failed) memcpy(w->cex, base[j].b, 8);
            w->failed = 1;
            pthread_mutex_unlock(&w->mu);
            return NULL;
          }
      }
    }
  }
  return NULL;
}

int sxo_exhaustive_brute(const uint32_t *t, int tn, const uint32_t *c, int cn, uint8_t dead, int nthreads,
                         uint8_t cex[8]) {
  brute_ctx w;
  memset(&w, 0, sizeof(w));
  w.t = t;
  w.c = c;
  w.tn = tn;
  w.cn = cn;
  w.dead = dead;
  const int dom = domain_bytes(t, tn, c, cn);
  for (int r = 0; r < 8; r++)
    if (r != SXO_A && r != SXO_P && (dom & (1 << r))) w.regs[w.k++] = r;
  w.nouter = 1ull << (8 * w.k);
  pthread_mutex_init(&w.mu, NULL);
  int nth = nthreads > 0 ? nthreads : 1;
  pthread_t th[64];
  if (nth > 64) nth = 64;
  for (int i = 0; i < nth; i++) pthread_create(&th[i], NULL, brute_worker, &w);
  for (int i = 0; i < nth; i++) pthread_join(th[i], NULL);
  pthread_mutex_destroy(&w.mu);
  if (w.failed && cex) memcpy(cex, w.cex, 8);
  return !w.failed;
}

/* Dependency units: state byte r (A=0, X=1, Y=2, S=4, M=5, S0=6, S1=7) and
 * bit b of P as unit 8 + b.  sxo_dep_units(op, d): d[u] = the input units
 * that output unit u of opcode op is a function of (an over-approximation),
 * d[u] = 1 << u where op leaves u unchanged; d[3] (the P byte) = 0.  Written
 * from the executor above (exec.go:8-327), one line per opcode family. */
#define UA (1u << SXO_A)
#define UX (1u << SXO_X)
#define UY (1u << SXO_Y)
#define US (1u << SXO_S)
#define UM (1u << SXO_M)
#define US0 (1u << SXO_S0)
#define US1 (1u << SXO_S1)
#define UP(b) (1u << (8 + (b)))
#define UC UP(0)
#define UALLP 0xFF00u
int sxo_dep_units(int op, uint16_t d[16]) {
  if (op < 0 || op >= SXO_OPCOUNT) return -1;
  for (int u = 0; u < 16; u++) d[u] = (uint16_t)(u == 3 ? 0 : 1u << u);
  uint32_t nz = 0xFFFFFFFFu; /* N and Z come from this set (none when unchanged) */
#define SET(u, m) d[u] = (uint16_t)(m)
  switch (op) {
  case SXO_TAX: SET(SXO_X, UA); nz = UA; break;
  case SXO_TAY: SET(SXO_Y, UA); nz = UA; break;
  case SXO_TXA: SET(SXO_A, UX); nz = UX; break;
  case SXO_TYA: SET(SXO_A, UY); nz = UY; break;
  case SXO_TSX: SET(SXO_X, US); nz = US; break;
  case SXO_TXS: SET(SXO_S, UX); break;
  case SXO_INX: case SXO_DEX: SET(SXO_X, UX); nz = UX; break;
  case SXO_INY: case SXO_DEY: SET(SXO_Y, UY); nz = UY; break;
  case SXO_ASL_A: case SXO_LSR_A: SET(SXO_A, UA); SET(8, UA); nz = UA; break;
  case SXO_ROL_A: case SXO_ROR_A: SET(SXO_A, UA | UC); SET(8, UA); nz = UA | UC; break;
  case SXO_CLC: case SXO_SEC: SET(8, 0); break;
  case SXO_CLI: case SXO_SEI: SET(8 + 2, 0); break;
  case SXO_CLD: case SXO_SED: SET(8 + 3, 0); break;
  case SXO_CLV: SET(8 + 6, 0); break;
  case SXO_PHA: SET(SXO_S1, US0); SET(SXO_S0, UA); break;
  case SXO_PHP: SET(SXO_S1, US0); SET(SXO_S0, UALLP); break;
  case SXO_PLA: SET(SXO_A, US0); SET(SXO_S0, US1); SET(SXO_S1, 0); nz = US0; break;
  case SXO_PLP:
    for (int b = 0; b < 8; b++) SET(8 + b, US0);
    SET(SXO_S0, US1); SET(SXO_S1, 0);
    break;
  case SXO_NOP: break;
  case SXO_LDA_IMM: SET(SXO_A, 0); nz = 0; break;
  case SXO_LDX_IMM: SET(SXO_X, 0); nz = 0; break;
  case SXO_LDY_IMM: SET(SXO_Y, 0); nz = 0; break;
  case SXO_ADC_IMM: case SXO_SBC_IMM:
    SET(SXO_A, UA | UC); SET(8, UA | UC); SET(8 + 6, UA | UC); nz = UA | UC; break;
  case SXO_AND_IMM: case SXO_ORA_IMM: case SXO_EOR_IMM: SET(SXO_A, UA); nz = UA; break;
  case SXO_CMP_IMM: SET(8, UA); nz = UA; break;
  case SXO_CPX_IMM: SET(8, UX); nz = UX; break;
  case SXO_CPY_IMM: SET(8, UY); nz = UY; break;
  case SXO_LDA_M: SET(SXO_A, UM); nz = UM; break;
  case SXO_LDX_M: SET(SXO_X, UM); nz = UM; break;
  case SXO_LDY_M: SET(SXO_Y, UM); nz = UM; break;
  case SXO_STA_M: SET(SXO_M, UA); break;
  case SXO_STX_M: SET(SXO_M, UX); break;
  case SXO_STY_M: SET(SXO_M, UY); break;
  case SXO_ADC_M: case SXO_SBC_M:
    SET(SXO_A, UA | UM | UC); SET(8, UA | UM | UC); SET(8 + 6, UA | UM | UC); nz = UA | UM | UC; break;
  case SXO_AND_M: case SXO_ORA_M: case SXO_EOR_M: SET(SXO_A, UA | UM); nz = UA | UM; break;
  case SXO_CMP_M: SET(8, UA | UM); nz = UA | UM; break;
  case SXO_CPX_M: SET(8, UX | UM); nz = UX | UM; break;
  case SXO_CPY_M: SET(8, UY | UM); nz = UY | UM; break;
  case SXO_INC_M: case SXO_DEC_M: SET(SXO_M, UM); nz = UM; break;
  case SXO_ASL_M: case SXO_LSR_M: SET(SXO_M, UM); SET(8, UM); nz = UM; break;
  case SXO_ROL_M: case SXO_ROR_M: SET(SXO_M, UM | UC); SET(8, UM); nz = UM | UC; break;
  case SXO_BIT_M: SET(8 + 1, UA | UM); SET(8 + 6, UM); SET(8 + 7, UM); break;
  }
  if (nz != 0xFFFFFFFFu) {
    SET(8 + 1, nz);
    SET(8 + 7, nz);
  }
#undef SET
  return 0;
}

/* Output units of a sequence as functions of its input units: compose the
 * per-opcode table along the sequence (out[u] = input units u depends on). */
static void seq_deps(const uint32_t *seq, int n, uint32_t out[16]) {
  for (int u = 0; u < 16; u++) out[u] = u == 3 ? 0 : 1u << u;
  for (int i = 0; i < n; i++) {
    uint16_t d[16];
    sxo_dep_units(SXO_OP(seq[i]), d);
    uint32_t nd[16];
    for (int u = 0; u < 16; u++) {
      nd[u] = 0;
      for (int q = 0; q < 16; q++)
        if (d[u] & (1u << q)) nd[u] |= out[q];
    }
    memcpy(out, nd, sizeof(nd));
  }
}

/* verifier.go:262-415 (dead == 0) and :457-588 (dead != 0).  Returns 1 when
 * the pair passes, else 0 with a state of the sweep on which the outputs
 * differ (cex may be NULL).
 *   mode 0: Go: every value of up to two extra registers
 *           (exhaustiveFullSweep), 32 representative values each for three or
 *           more (exhaustiveReducedSweep);
 *   mode 1: complete, by brute force over the whole domain (2^(9+8k) states);
 *   mode 2: complete, by projection: every compared output unit (a state byte
 *           or a flag bit of P under PMask &^ dead) of T and of C is a
 *           function of the inputs sxo_dep_units composes for it, so T and C
 *           agree on it over the whole domain iff they agree on the sweep of
 *           the domain bytes among those inputs (the other bytes at their
 *           defaults).  Each such sweep compares the whole masked state, so a
 *           disagreement there is a real counterexample in the domain.
 * The visiting order differs from Go's nested loops, which does not change
 * the boolean result. */
int sxo_exhaustive_cex(const uint32_t *t, int tn, const uint32_t *c, int cn, uint8_t dead, int mode,
                       uint8_t cex[8]) {
  const int dom = domain_bytes(t, tn, c, cn);
  if (mode == 1) return sweep_set(t, tn, c, cn, dead, dom, cex);
  if (mode == 2) {
    uint32_t dt[16], dc[16];
    seq_deps(t, tn, dt);
    seq_deps(c, cn, dc);
    /* domain bytes as input units: the carry is unit 8 */
    const uint32_t dom_units = (uint32_t)(dom & ~(1 << SXO_P)) | UC;
    int done[16], nd = 0;
    for (int u = 0; u < 16; u++) {
      if (u == 3 || (u >= 8 && !((PMASK & ~dead) & (1 << (u - 8))))) continue; /* not compared */
      const uint32_t in = (dt[u] | dc[u]) & dom_units;
      const int set = (int)(in & 0xF7u) | ((in & UC) ? 1 << SXO_P : 0);
      int seen = 0;
      for (int i = 0; i < nd; i++) seen |= done[i] == set;
      if (seen) continue;
      done[nd++] = set;
      if (!sweep_set(t, tn, c, cn, dead, set, cex)) return 0;
    }
    return 1;
  }
  /* mode 0 */
  int reads = sxo_regs_read(t, tn) | sxo_regs_read(c, cn);
  int ex[6];
  int k = (reads & ~(SXO_RA | SXO_RP)) ? sxo_extra_regs(reads, ex) : 0;
  if (k < 3) return sweep_set(t, tn, c, cn, dead, dom, cex);
  int eb[6];
  for (int i = 0; i < k; i++) eb[i] = k_off_to_byte[ex[i]];
  int odo[6] = {0, 0, 0, 0, 0, 0};
  st8 base;
  base.q = 0;
  base.b[SXO_S] = 0xFD;
  for (;;) {
    for (int i = 0; i < k; i++) base.b[eb[i]] = k_rep[odo[i]];
    for (int carry = 0; carry <= 1; carry++) {
      base.b[SXO_P] = (uint8_t)carry;
      for (int a = 0; a < 256; a++) {
        base.b[SXO_A] = (uint8_t)a;
        if (!pair_agrees(&base, t, tn, c, cn, dead)) {
          if (cex) memcpy(cex, base.b, 8);
          return 0;
        }
      }
    }
    int i = 0;
    while (i < k && ++odo[i] == 32) odo[i++] = 0;
    if (i == k) break;
  }
  return 1;
}

int sxo_exhaustive_check(const uint32_t *t, int tn, const uint32_t *c, int cn, uint8_t dead, int mode) {
  return sxo_exhaustive_cex(t, tn, c, cn, dead, mode, NULL);
}

/* ------------------------------------------------------ Pruner / enum -- */

/* pruner.go:38-53 */
static int is_dead_write(uint32_t first, uint32_t second) {
  int w = sxo_op_writes(SXO_OP(first));
  if (!w) return 0;
  int r = sxo_op_reads(SXO_OP(second));
  int w2 = sxo_op_writes(SXO_OP(second));
  int dead = w & w2 & ~SXO_RP & ~SXO_RS & ~SXO_RS0 & ~SXO_RS1 & ~r;
  return dead != 0;
}

/* pruner.go:57-75 */
static int has_stack_violation(const uint32_t *seq, int n) {
  int depth = 0;
  for (int i = 0; i < n; i++) {
    int op = SXO_OP(seq[i]);
    if (op == SXO_PHA || op == SXO_PHP) {
      if (++depth > 2) return 1;
    } else if (op == SXO_PLA || op == SXO_PLP) {
      if (depth <= 0) return 1;
      depth--;
    }
  }
  return 0;
}

/* pruner.go:166-181 */
static int are_independent(uint32_t a, uint32_t b) {
  int aR = sxo_op_reads(SXO_OP(a)), aW = sxo_op_writes(SXO_OP(a));
  int bR = sxo_op_reads(SXO_OP(b)), bW = sxo_op_writes(SXO_OP(b));
  return !((aW & bR) || (aR & bW) || (aW & bW));
}

/* pruner.go:6-34; instKey = op<<16 | imm is the instruction word itself. */
int sxo_should_prune(const uint32_t *seq, int n) {
  for (int i = 0; i < n; i++) {
    if (SXO_OP(seq[i]) == SXO_NOP) return 1;
    if (i + 1 < n && is_dead_write(seq[i], seq[i + 1])) return 1;
  }
  if (has_stack_violation(seq, n)) return 1;
  for (int i = 0; i + 1 < n; i++) {
    uint32_t ka = ((uint32_t)SXO_OP(seq[i]) << 16) | SXO_IMM(seq[i]);
    uint32_t kb = ((uint32_t)SXO_OP(seq[i + 1]) << 16) | SXO_IMM(seq[i + 1]);
    if (are_independent(seq[i], seq[i + 1]) && ka > kb) return 1;
  }
  return 0;
}

/* search.go:56-75 with enumerator.go:7-38 (lexicographic over Go order). */
int64_t sxo_enumerate_pruned(int len, uint32_t *out, int64_t cap) {
  uint32_t tab[2863];
  int n = sxo_instruction_table(tab, 0);
  int64_t cnt = 0;
  uint32_t seq[3];
  if (len == 1) {
    for (int i = 0; i < n; i++) {
      seq[0] = tab[i];
      if (sxo_should_prune(seq, 1)) continue;
      if (out && cnt < cap) out[cnt] = seq[0];
      cnt++;
    }
  } else if (len == 2) {
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        seq[0] = tab[i];
        seq[1] = tab[j];
        if (sxo_should_prune(seq, 2)) continue;
        if (out && cnt < cap) memcpy(out + 2 * cnt, seq, 8);
        cnt++;
      }
  } else {
    return -1;
  }
  return cnt;
}

/* -------------------------------------------------------- Search driver -- */

typedef struct cand_info {
  uint32_t w[3];
  int len;
  int bytes;
  int cycles;
  int prune;
} cand_info;

typedef struct worker_ctx {
  const sxo_search_cfg *cfg;
  const uint32_t *targets;
  int64_t ntargets;
  const cand_info *cands[3]; /* by length (1, 2; length 3 is enumerated on the fly) */
  int64_t ncands[3];
  uint32_t tab3[2863];       /* length 3: the instructions in Go order and their sizes */
  int bytes3[2863];
  int n3;
  int64_t next;              /* shared cursor, atomic */
  pthread_mutex_t mu;
  sxo_rule *rules;
  int64_t nrules, caprules;
  sxo_stats stats;
} worker_ctx;

static void push_rule(worker_ctx *w, sxo_rule *r) {
  pthread_mutex_lock(&w->mu);
  if (w->nrules == w->caprules) {
    w->caprules = w->caprules ? 2 * w->caprules : 4096;
    w->rules = (sxo_rule *)realloc(w->rules, (size_t)w->caprules * sizeof(sxo_rule));
  }
  w->rules[w->nrules++] = *r;
  pthread_mutex_unlock(&w->mu);
}

/* QuickCheck against precomputed target outputs on the 8 vectors: the same
 * predicate as verifier.go:84-93, evaluated without re-running the target
 * for every candidate. */
static int qc_pre(const uint8_t tout[8][8], const uint32_t *c, int cn, uint8_t dead) {
  for (int i = 0; i < 8; i++) {
    uint8_t cs[8];
    memcpy(cs, k_qc[i], 8);
    sxo_exec_seq(cs, c, cn);
    if (!sxo_states_equal_masked(tout[i], cs, dead)) return 0;
  }
  return 1;
}

static void make_rule(sxo_rule *r, const uint32_t *t, int tn, const cand_info *ci,
                      int64_t ti, int tbytes, int tcycles, uint8_t dead, int pass) {
  memset(r, 0, sizeof(*r));
  for (int i = 0; i < tn; i++) r->src[i] = t[i];
  for (int i = 0; i < ci->len; i++) r->rep[i] = ci->w[i];
  r->src_len = (uint8_t)tn;
  r->rep_len = (uint8_t)ci->len;
  r->dead_flags = dead;
  r->pass = (uint8_t)pass;
  r->bytes_saved = (int16_t)(tbytes - ci->bytes);
  r->cycles_saved = (int16_t)(tcycles - ci->cycles);
  r->target_index = (uint64_t)ti;
}

/* Length-3 candidates of one pass, in Go order (EnumerateSequences(3),
 * enumerator.go: first instruction outermost), generated on the fly (there
 * are 2,863^3): the first candidate that passes the byte filter, the pruner,
 * the pass's QuickCheck, [MidCheck], the ExhaustiveCheck and, in the
 * flag-relaxed pass, FlagDiff != 0 (worker.go:122-170, 180-240).  pass 1:
 * the exact QuickCheck; pass 2: fails the exact QuickCheck and passes the
 * masked one.  The 8 vectors' states after the first two instructions are
 * computed once per prefix, and the QuickCheck looks at vector 0 first; the
 * filters are conjunctive, so their order does not change the first hit.
 * checked counts every candidate enumerated before the loop stops, as
 * worker.go:126 does.  Returns 1 with the candidate in *out when found. */
static int scan_len3(const worker_ctx *w, const uint32_t *t, int tn, int tbytes, const uint8_t tout[8][8],
                     int pass, uint8_t dead, cand_info *out, sxo_stats *st) {
  const sxo_search_cfg *cfg = w->cfg;
  const uint32_t *tab = w->tab3;
  const int *bm = w->bytes3;
  const int n = w->n3;
  const uint8_t d_qc = pass == 1 ? 0 : dead;
  /* the byte filter in bulk: suf[i] = the fewest bytes of instructions i..n-1
   * (Go order puts the one-byte instructions first), so once
   * pb + suf[i2] >= tbytes no later last instruction fits either; skipped
   * candidates still count as checked */
  int suf[2864];
  suf[n] = 1 << 20;
  for (int i = n - 1; i >= 0; i--) suf[i] = bm[i] < suf[i + 1] ? bm[i] : suf[i + 1];
  const uint64_t nn = (uint64_t)n * (uint64_t)n;
  for (int i0 = 0; i0 < n; i0++) {
    if (bm[i0] + suf[0] + suf[0] >= tbytes) {
      st->checked += nn;
      continue;
    }
    for (int i1 = 0; i1 < n; i1++) {
      const int pb = bm[i0] + bm[i1];
      uint8_t ps[8][8];
      int have = 0;
      if (pb + suf[0] >= tbytes) {
        st->checked += (uint64_t)n;
        continue;
      }
      for (int i2 = 0; i2 < n; i2++) {
        if (pb + suf[i2] >= tbytes) {
          st->checked += (uint64_t)(n - i2);
          break;
        }
        st->checked++;
        if (pb + bm[i2] >= tbytes) continue;
        if (!have) {
          const uint32_t pre[2] = {tab[i0], tab[i1]};
          for (int v = 0; v < 8; v++) {
            memcpy(ps[v], k_qc[v], 8);
            sxo_exec_seq(ps[v], pre, 2);
          }
          have = 1;
        }
        uint8_t s0[8];
        memcpy(s0, ps[0], 8);
        exec_one(s0, SXO_OP(tab[i2]), SXO_IMM(tab[i2]));
        if (!sxo_states_equal_masked(tout[0], s0, d_qc)) continue;
        int ok = 1, exact = sxo_states_equal_masked(tout[0], s0, 0);
        for (int v = 1; v < 8 && ok; v++) {
          uint8_t sv[8];
          memcpy(sv, ps[v], 8);
          exec_one(sv, SXO_OP(tab[i2]), SXO_IMM(tab[i2]));
          ok = sxo_states_equal_masked(tout[v], sv, d_qc);
          exact = exact && sxo_states_equal_masked(tout[v], sv, 0);
        }
        if (!ok || (pass == 2 && exact)) continue;
        const uint32_t c[3] = {tab[i0], tab[i1], tab[i2]};
        if (sxo_should_prune(c, 3)) continue;
        st->qc_hits++;
        if (cfg->use_mid && !sxo_mid_check(t, tn, c, 3, d_qc)) continue;
        st->exh_calls++;
        if (!sxo_exhaustive_check(t, tn, c, 3, d_qc, cfg->exh_mode)) continue;
        if (pass == 2 && sxo_flag_diff(t, tn, c, 3) == 0) continue;
        memset(out, 0, sizeof(*out));
        out->w[0] = c[0];
        out->w[1] = c[1];
        out->w[2] = c[2];
        out->len = 3;
        out->bytes = pb + bm[i2];
        for (int k = 0; k < 3; k++) out->cycles += sxo_cycles(SXO_OP(c[k]));
        return 1;
      }
    }
  }
  return 0;
}

static void process_target(worker_ctx *w, int64_t ti, sxo_stats *st) {
  const sxo_search_cfg *cfg = w->cfg;
  int tn = cfg->target_len;
  const uint32_t *t = w->targets + ti * tn;
  int tbytes = 0, tcycles = 0;
  for (int i = 0; i < tn; i++) {
    tbytes += sxo_byte_size(SXO_OP(t[i]), cfg->bytes_model);
    tcycles += sxo_cycles(SXO_OP(t[i]));
  }
  uint8_t tout[8][8];
  for (int i = 0; i < 8; i++) {
    memcpy(tout[i], k_qc[i], 8);
    sxo_exec_seq(tout[i], t, tn);
  }
  sxo_rule r;
  if (cfg->include_empty) {
    /* The empty sequence as a candidate, in a stream of its own: Go never
     * tries it (worker.go:123 starts at length 1), so it does not take part
     * in the first-hit order of the other candidates.  It has 0 bytes and
     * is never pruned; pass 1 is the exact check, pass 2 (dead flags) the
     * flag-relaxed one with FlagDiff != 0; in the CUDA tool's semantics one
     * masked pass. */
    static const cand_info empty = {{0, 0}, 0, 0, 0, 0};
    const uint8_t d1 = cfg->hits == 1 ? cfg->dead : 0;
    st->checked++;
    if (qc_pre(tout, empty.w, 0, d1) && (!(cfg->use_mid || cfg->hits == 1) || sxo_mid_check(t, tn, empty.w, 0, d1)) &&
        sxo_exhaustive_check(t, tn, empty.w, 0, d1, cfg->exh_mode)) {
      make_rule(&r, t, tn, &empty, ti, tbytes, tcycles, d1, 1);
      push_rule(w, &r);
    } else if (cfg->hits == 0 && cfg->dead && !qc_pre(tout, empty.w, 0, 0) && qc_pre(tout, empty.w, 0, cfg->dead) &&
               (!cfg->use_mid || sxo_mid_check(t, tn, empty.w, 0, cfg->dead)) &&
               sxo_exhaustive_check(t, tn, empty.w, 0, cfg->dead, cfg->exh_mode)) {
      const uint8_t fd = sxo_flag_diff(t, tn, empty.w, 0);
      if (fd) {
        make_rule(&r, t, tn, &empty, ti, tbytes, tcycles, fd, 2);
        push_rule(w, &r);
      }
    }
  }
  if (cfg->hits == 1) {
    /* CUDA tool: one masked pass, every hit (cuda/6502_search_v2.cu:871-974) */
    for (int len = 1; len <= cfg->max_cand_len; len++) {
      for (int64_t j = 0; j < w->ncands[len]; j++) {
        const cand_info *ci = &w->cands[len][j];
        st->checked++;
        if (ci->bytes >= tbytes || ci->prune) continue;
        if (!qc_pre(tout, ci->w, ci->len, cfg->dead)) continue;
        st->qc_hits++;
        if (!sxo_mid_check(t, tn, ci->w, ci->len, cfg->dead)) continue;
        st->mid_pass++;
        st->exh_calls++;
        if (!sxo_exhaustive_check(t, tn, ci->w, ci->len, cfg->dead, cfg->exh_mode)) continue;
        st->found++;
        make_rule(&r, t, tn, ci, ti, tbytes, tcycles, cfg->dead, 1);
        push_rule(w, &r);
      }
    }
    return;
  }
  /* worker.go:119-178: first hit per target, candidate lengths ascending. */
  for (int len = 1; len <= cfg->max_cand_len; len++) {
    int found = 0;
    if (len == 3) {
      cand_info c3;
      if (scan_len3(w, t, tn, tbytes, tout, 1, 0, &c3, st)) {
        st->found++;
        make_rule(&r, t, tn, &c3, ti, tbytes, tcycles, 0, 1);
        push_rule(w, &r);
      }
      break;
    }
    for (int64_t j = 0; j < w->ncands[len]; j++) {
      const cand_info *ci = &w->cands[len][j];
      st->checked++;
      if (ci->bytes >= tbytes) continue;
      if (ci->prune) continue;
      if (!qc_pre(tout, ci->w, ci->len, 0)) continue;
      st->qc_hits++;
      if (cfg->use_mid && !sxo_mid_check(t, tn, ci->w, ci->len, 0)) continue;
      st->exh_calls++;
      if (!sxo_exhaustive_check(t, tn, ci->w, ci->len, 0, cfg->exh_mode)) continue;
      st->found++;
      make_rule(&r, t, tn, ci, ti, tbytes, tcycles, 0, 1);
      push_rule(w, &r);
      found = 1;
      break;
    }
    if (found) break;
  }
  if (cfg->dead == 0) return;
  /* worker.go:180-245: flag-relaxed pass over candidates that fail the
   * exact QuickCheck. */
  for (int len = 1; len <= cfg->max_cand_len; len++) {
    int found = 0;
    if (len == 3) {
      cand_info c3;
      if (scan_len3(w, t, tn, tbytes, tout, 2, cfg->dead, &c3, st)) {
        st->found++;
        make_rule(&r, t, tn, &c3, ti, tbytes, tcycles, sxo_flag_diff(t, tn, c3.w, 3), 2);
        push_rule(w, &r);
      }
      break;
    }
    for (int64_t j = 0; j < w->ncands[len]; j++) {
      const cand_info *ci = &w->cands[len][j];
      st->checked++;
      if (ci->bytes >= tbytes) continue;
      if (ci->prune) continue;
      if (qc_pre(tout, ci->w, ci->len, 0)) continue;
      if (!qc_pre(tout, ci->w, ci->len, cfg->dead)) continue;
      st->qc_hits++;
      if (cfg->use_mid && !sxo_mid_check(t, tn, ci->w, ci->len, cfg->dead)) continue;
      st->exh_calls++;
      if (!sxo_exhaustive_check(t, tn, ci->w, ci->len, cfg->dead, cfg->exh_mode)) continue;
      uint8_t fd = sxo_flag_diff(t, tn, ci->w, ci->len);
      if (fd == 0) continue;
      st->found++;
      make_rule(&r, t, tn, ci, ti, tbytes, tcycles, fd, 2);
      push_rule(w, &r);
      found = 1;
      break;
    }
    if (found) break;
  }
}

static void *worker_main(void *arg) {
  worker_ctx *w = (worker_ctx *)arg;
  sxo_stats st;
  memset(&st, 0, sizeof(st));
  /* one target at a time when a target can take minutes (length-3 candidates) */
  const int64_t chunk = w->cfg->max_cand_len == 3 ? 1 : 64;
  for (;;) {
    int64_t b = __atomic_fetch_add(&w->next, chunk, __ATOMIC_RELAXED);
    if (b >= w->ntargets) break;
    int64_t e = b + chunk < w->ntargets ? b + chunk : w->ntargets;
    for (int64_t ti = b; ti < e; ti++) process_target(w, ti, &st);
  }
  pthread_mutex_lock(&w->mu);
  w->stats.checked += st.checked;
  w->stats.qc_hits += st.qc_hits;
  w->stats.mid_pass += st.mid_pass;
  w->stats.exh_calls += st.exh_calls;
  w->stats.found += st.found;
  pthread_mutex_unlock(&w->mu);
  return NULL;
}

static int rule_cmp(const void *a, const void *b) {
  const sxo_rule *x = (const sxo_rule *)a, *y = (const sxo_rule *)b;
  if (x->target_index != y->target_index) return x->target_index < y->target_index ? -1 : 1;
  if (x->pass != y->pass) return x->pass < y->pass ? -1 : 1;
  if (x->rep_len != y->rep_len) return x->rep_len < y->rep_len ? -1 : 1;
  for (int i = 0; i < 4; i++)
    if (x->rep[i] != y->rep[i]) return x->rep[i] < y->rep[i] ? -1 : 1;
  return 0;
}

static cand_info *build_cands(int len, int bytes_model, int64_t *count) {
  uint32_t tab[2863];
  int n = sxo_instruction_table(tab, 0);
  int64_t total = len == 1 ? n : (int64_t)n * n;
  cand_info *c = (cand_info *)calloc((size_t)total, sizeof(cand_info));
  int64_t k = 0;
  if (len == 1) {
    for (int i = 0; i < n; i++, k++) {
      c[k].w[0] = tab[i];
      c[k].len = 1;
    }
  } else {
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++, k++) {
        c[k].w[0] = tab[i];
        c[k].w[1] = tab[j];
        c[k].len = 2;
      }
  }
  for (int64_t i = 0; i < total; i++) {
    for (int p = 0; p < c[i].len; p++) {
      c[i].bytes += sxo_byte_size(SXO_OP(c[i].w[p]), bytes_model);
      c[i].cycles += sxo_cycles(SXO_OP(c[i].w[p]));
    }
    c[i].prune = sxo_should_prune(c[i].w, c[i].len);
  }
  *count = total;
  return c;
}

int sxo_search(const sxo_search_cfg *cfg, const uint32_t *targets, int64_t ntargets,
               sxo_rule *out, int64_t cap, int64_t *nout, sxo_stats *stats) {
  if (!cfg || cfg->target_len < 1 || cfg->target_len > 4 || cfg->max_cand_len < 1 ||
      cfg->max_cand_len > 3 || (cfg->max_cand_len == 3 && cfg->hits))
    return -1;
  worker_ctx w;
  memset(&w, 0, sizeof(w));
  w.cfg = cfg;
  w.targets = targets;
  w.ntargets = ntargets;
  pthread_mutex_init(&w.mu, NULL);
  for (int len = 1; len <= cfg->max_cand_len && len <= 2; len++)
    w.cands[len] = build_cands(len, cfg->bytes_model, &w.ncands[len]);
  if (cfg->max_cand_len == 3) {
    w.n3 = sxo_instruction_table(w.tab3, 0);
    for (int i = 0; i < w.n3; i++) w.bytes3[i] = sxo_byte_size(SXO_OP(w.tab3[i]), cfg->bytes_model);
  }
  int nt = cfg->nthreads > 0 ? cfg->nthreads : 1;
  pthread_t *th = (pthread_t *)calloc((size_t)nt, sizeof(pthread_t));
  for (int i = 0; i < nt; i++) pthread_create(&th[i], NULL, worker_main, &w);
  for (int i = 0; i < nt; i++) pthread_join(th[i], NULL);
  free(th);
  for (int len = 1; len <= cfg->max_cand_len && len <= 2; len++) free((void *)w.cands[len]);
  qsort(w.rules, (size_t)w.nrules, sizeof(sxo_rule), rule_cmp);
  if (nout) *nout = w.nrules;
  if (out) {
    int64_t m = w.nrules < cap ? w.nrules : cap;
    if (m > 0) memcpy(out, w.rules, (size_t)m * sizeof(sxo_rule));
  }
  if (stats) *stats = w.stats;
  free(w.rules);
  pthread_mutex_destroy(&w.mu);
  return (out && w.nrules > cap) ? -3 : 0;
}

/* ------------------------------------------- CUDA-tool fallback pairs -- */

typedef struct fb_ctx {
  const uint32_t *targets; /* pruned length-2 targets, Go order */
  int64_t ntargets;
  const uint32_t *cands;   /* 2,863 single instructions, Go order */
  uint8_t cout[2863][32][8];
  int creads[2863], cbytes[2863], cprune[2863];
  int min_extra;
  int64_t next;
  pthread_mutex_t mu;
  uint32_t *pairs; /* (target index, candidate index) */
  int64_t npairs, cap;
} fb_ctx;

static int extra_count(int reads) {
  return __builtin_popcount((unsigned)reads & (SXO_RX | SXO_RY | SXO_RS | SXO_RM | SXO_RS0 | SXO_RS1));
}

static void *fb_worker(void *arg) {
  fb_ctx *w = (fb_ctx *)arg;
  const int64_t chunk = 4096;
  uint32_t *loc = NULL;
  int64_t nloc = 0, caploc = 0;
  for (;;) {
    int64_t b = __atomic_fetch_add(&w->next, chunk, __ATOMIC_RELAXED);
    if (b >= w->ntargets) break;
    int64_t e = b + chunk < w->ntargets ? b + chunk : w->ntargets;
    for (int64_t ti = b; ti < e; ti++) {
      const uint32_t *t = w->targets + 2 * ti;
      const int treads = sxo_regs_read(t, 2);
      const int tbytes = sxo_byte_size(SXO_OP(t[0]), 1) + sxo_byte_size(SXO_OP(t[1]), 1);
      uint8_t tout[32][8];
      int have_out = 0;
      for (int j = 0; j < 2863; j++) {
        if (extra_count(treads | w->creads[j]) < w->min_extra) continue;
        if (w->cbytes[j] >= tbytes || w->cprune[j]) continue; /* cu:877-880 */
        if (!have_out) {
          for (int v = 0; v < 32; v++) {
            memcpy(tout[v], k_mid[v], 8);
            sxo_exec_seq(tout[v], t, 2);
          }
          have_out = 1;
        }
        int ok = 1;
        for (int v = 0; v < 32 && ok; v++) ok = sxo_states_equal_masked(tout[v], w->cout[j][v], 0);
        if (!ok) continue;
        if (nloc + 2 > caploc) {
          caploc = caploc ? 2 * caploc : 4096;
          loc = (uint32_t *)realloc(loc, (size_t)caploc * sizeof(uint32_t));
        }
        loc[nloc++] = (uint32_t)ti;
        loc[nloc++] = (uint32_t)j;
      }
    }
  }
  pthread_mutex_lock(&w->mu);
  if (w->npairs + nloc > w->cap) {
    w->cap = (w->npairs + nloc) * 2;
    w->pairs = (uint32_t *)realloc(w->pairs, (size_t)w->cap * sizeof(uint32_t));
  }
  if (nloc) memcpy(w->pairs + w->npairs, loc, (size_t)nloc * sizeof(uint32_t));
  w->npairs += nloc;
  pthread_mutex_unlock(&w->mu);
  free(loc);
  return NULL;
}

static int pair_cmp(const void *a, const void *b) {
  const uint32_t *x = (const uint32_t *)a, *y = (const uint32_t *)b;
  if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
  return x[1] < y[1] ? -1 : (x[1] > y[1]);
}

int64_t sxo_cuda_fallback_pairs(int min_extra, uint32_t *out, int64_t cap, int nthreads) {
  int64_t nt = sxo_enumerate_pruned(2, NULL, 0);
  uint32_t *targets = (uint32_t *)malloc((size_t)nt * 2 * sizeof(uint32_t));
  sxo_enumerate_pruned(2, targets, nt);
  uint32_t cands[2863];
  sxo_instruction_table(cands, 0);
  fb_ctx *w = (fb_ctx *)calloc(1, sizeof(fb_ctx));
  w->targets = targets;
  w->ntargets = nt;
  w->cands = cands;
  w->min_extra = min_extra;
  for (int j = 0; j < 2863; j++) {
    w->creads[j] = sxo_op_reads(SXO_OP(cands[j]));
    w->cbytes[j] = sxo_byte_size(SXO_OP(cands[j]), 1);
    w->cprune[j] = sxo_should_prune(&cands[j], 1);
    for (int v = 0; v < 32; v++) {
      memcpy(w->cout[j][v], k_mid[v], 8);
      sxo_exec_seq(w->cout[j][v], &cands[j], 1);
    }
  }
  pthread_mutex_init(&w->mu, NULL);
  int nth = nthreads > 0 ? nthreads : 1;
  pthread_t *th = (pthread_t *)calloc((size_t)nth, sizeof(pthread_t));
  for (int i = 0; i < nth; i++) pthread_create(&th[i], NULL, fb_worker, w);
  for (int i = 0; i < nth; i++) pthread_join(th[i], NULL);
  free(th);
  const int64_t n = w->npairs / 2;
  qsort(w->pairs, (size_t)n, 2 * sizeof(uint32_t), pair_cmp);
  if (out)
    for (int64_t i = 0; i < n && i < cap; i++) {
      out[3 * i] = targets[2 * (int64_t)w->pairs[2 * i]];
      out[3 * i + 1] = targets[2 * (int64_t)w->pairs[2 * i] + 1];
      out[3 * i + 2] = cands[w->pairs[2 * i + 1]];
    }
  pthread_mutex_destroy(&w->mu);
  free(w->pairs);
  free(w);
  free(targets);
  return n;
}

/* ------------------------------------------------------------------ STOKE -- */

/* Philox4x32-10: 10 rounds of the S-box (two 32x32->64 multiplies) with the
 * Weyl key schedule; multipliers and key increments from the SC'11 paper. */
void sxo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static int seq_bytes_go(const uint32_t *s, int n) {
  int b = 0;
  for (int i = 0; i < n; i++) b += sxo_byte_size(SXO_OP(s[i]), 0);
  return b;
}

/* Hamming distance of the output states over the 32 Mid vectors, P compared
 * under 0xCF &^ dead (pkg/stoke/cost.go:32-81 counts mismatching vectors on
 * the 8 QuickCheck vectors instead; this build's cost is defined by bits). */
static int hamming32(const uint8_t tout[32][8], const uint32_t *c, int cn, uint8_t dead) {
  int h = 0;
  uint8_t pm = (uint8_t)(PMASK & ~dead);
  for (int v = 0; v < 32; v++) {
    uint8_t s[8];
    memcpy(s, k_mid[v], 8);
    sxo_exec_seq(s, c, cn);
    for (int r = 0; r < 8; r++) {
      uint8_t d = (uint8_t)(s[r] ^ tout[v][r]);
      if (r == SXO_P) d &= pm;
      h += __builtin_popcount(d);
    }
  }
  return h;
}

static void target_outs(const uint32_t *t, int tn, uint8_t out[32][8]) {
  for (int v = 0; v < 32; v++) {
    memcpy(out[v], k_mid[v], 8);
    sxo_exec_seq(out[v], t, tn);
  }
}

int sxo_stoke_cost10(const uint32_t *t, int tn, const uint32_t *c, int cn, uint8_t dead) {
  uint8_t tout[32][8];
  target_outs(t, tn, tout);
  return 10 * hamming32(tout, c, cn, dead) + seq_bytes_go(c, cn);
}

void sxo_stoke_init(const uint32_t *t, int tn, double t0, uint8_t dead, sxo_chain *ch) {
  memset(ch, 0, sizeof(*ch));
  for (int i = 0; i < tn; i++) ch->seq[i] = t[i];
  ch->len = tn;
  ch->cost10 = sxo_stoke_cost10(t, tn, t, tn, dead);
  ch->temp = t0;
  ch->first_hit_step = -1;
}

/* Move-type thresholds on the top byte of the first random word: replace
 * 102/256, swap 52/256, delete 51/256, insert 25/256, immediate 26/256 (the
 * 40/20/20/10/10 split of pkg/stoke/mutator.go:30-44 in 1/256 units). */
#define MV_REPLACE_END 102u
#define MV_SWAP_END 154u
#define MV_DELETE_END 205u
#define MV_INSERT_END 230u
#define W_REP (102.0 / 256.0)
#define W_DEL (51.0 / 256.0)
#define W_INS (25.0 / 256.0)
#define W_IMM (26.0 / 256.0)

static inline uint32_t mulhi32(uint32_t a, uint32_t n) { return (uint32_t)(((uint64_t)a * n) >> 32); }

/* Instruction index with cumulative weight containing u (weights sum to 2^32). */
static int pick_inst(const uint32_t *w, uint32_t u) {
  uint64_t acc = 0;
  for (int i = 0; i < 2863; i++) {
    acc += w[i];
    if ((uint64_t)u < acc) return i;
  }
  return 2862;
}

static void uniform_weights(uint32_t w[2863]) {
  uint64_t total = (uint64_t)1 << 32;
  uint32_t base = (uint32_t)(total / 2863), extra = (uint32_t)(total % 2863);
  for (int i = 0; i < 2863; i++) w[i] = base + (i < (int)extra ? 1u : 0u);
}

void sxo_stoke_steps(const uint32_t *t, int tn, uint8_t dead, uint32_t chain_id, uint64_t seed,
                     double decay, const uint32_t *weights, int nsteps, sxo_chain *ch) {
  uint32_t tab[2863], uw[2863];
  sxo_instruction_table(tab, 0);
  if (!weights) {
    uniform_weights(uw);
    weights = uw;
  }
  /* Go-order index of an instruction word */
  int32_t idx_of[58][256];
  for (int i = 0; i < 2863; i++) idx_of[SXO_OP(tab[i])][SXO_IMM(tab[i]) & 0xFF] = i;
#define QW(word) ((double)weights[idx_of[SXO_OP(word)][sxo_has_imm(SXO_OP(word)) ? (SXO_IMM(word) & 0xFF) : 0]] / 4294967296.0)
  uint8_t tout[32][8];
  target_outs(t, tn, tout);
  const int tbytes = seq_bytes_go(t, tn);
  const int max_len = tn + 2 > 10 ? tn + 2 : 10;
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (int it = 0; it < nsteps; it++) {
    uint32_t ctr[4] = {(uint32_t)ch->step, (uint32_t)(ch->step >> 32), chain_id, 0}, r[4], a[4];
    sxo_philox4x32_10(ctr, key, r);
    ctr[3] = 1;
    sxo_philox4x32_10(ctr, key, a);
    uint32_t x[10];
    int L = ch->len, nl = L;
    memcpy(x, ch->seq, sizeof(x));
    uint32_t m = r[0] >> 24;
    double ratio = 1.0;
    int changed = 0;
    if (m < MV_REPLACE_END) {
      int pos = (int)mulhi32(r[1], (uint32_t)L);
      uint32_t v = tab[pick_inst(weights, r[2])];
      if (v != x[pos]) {
        double same = 0.0;
        int nimm = 0;
        for (int i = 0; i < L; i++) nimm += sxo_has_imm(SXO_OP(x[i]));
        if (SXO_OP(v) == SXO_OP(x[pos]) && sxo_has_imm(SXO_OP(v))) same = W_IMM / (nimm * 256.0);
        ratio = (W_REP * QW(x[pos]) / L + same) / (W_REP * QW(v) / L + same);
        x[pos] = v;
        changed = 1;
      }
    } else if (m < MV_SWAP_END) {
      if (L >= 2) {
        int pos = (int)mulhi32(r[1], (uint32_t)(L - 1));
        if (x[pos] != x[pos + 1]) {
          uint32_t tmp = x[pos];
          x[pos] = x[pos + 1];
          x[pos + 1] = tmp;
          changed = 1;
        }
      }
    } else if (m < MV_DELETE_END) {
      if (L > 1) {
        int pos = (int)mulhi32(r[1], (uint32_t)L);
        ratio = W_INS * QW(x[pos]) / W_DEL;
        for (int i = pos; i + 1 < L; i++) x[i] = x[i + 1];
        nl = L - 1;
        changed = 1;
      }
    } else if (m < MV_INSERT_END) {
      if (L < max_len) {
        int pos = (int)mulhi32(r[1], (uint32_t)(L + 1));
        uint32_t v = tab[pick_inst(weights, r[2])];
        ratio = W_DEL / (W_INS * QW(v));
        for (int i = L; i > pos; i--) x[i] = x[i - 1];
        x[pos] = v;
        nl = L + 1;
        changed = 1;
      }
    } else {
      int nimm = 0;
      for (int i = 0; i < L; i++) nimm += sxo_has_imm(SXO_OP(x[i]));
      if (nimm > 0) {
        int k = (int)mulhi32(r[1], (uint32_t)nimm), pos = 0;
        for (int i = 0; i < L; i++)
          if (sxo_has_imm(SXO_OP(x[i])) && k-- == 0) {
            pos = i;
            break;
          }
        uint32_t v = SXO_WORD(SXO_OP(x[pos]), r[3] >> 24);
        if (v != x[pos]) {
          double same = W_IMM / (nimm * 256.0);
          ratio = (W_REP * QW(x[pos]) / L + same) / (W_REP * QW(v) / L + same);
          x[pos] = v;
          changed = 1;
        }
      }
    }
    if (changed) {
      int h = hamming32(tout, x, nl, dead);
      int c10 = 10 * h + seq_bytes_go(x, nl);
      double la = -((double)(c10 - ch->cost10) / 10.0) / ch->temp + log(ratio);
      double u = ((double)(a[0] >> 5) * 67108864.0 + (double)(a[1] >> 6) + 0.5) / 9007199254740992.0;
      if (la >= 0.0 || u < exp(la)) {
        memcpy(ch->seq, x, sizeof(x));
        ch->len = nl;
        ch->cost10 = c10;
        ch->accepted++;
        if (h == 0 && seq_bytes_go(x, nl) < tbytes) {
          ch->hits++;
          if (ch->first_hit_step < 0) ch->first_hit_step = (int32_t)ch->step;
        }
      }
    }
    ch->temp *= decay;
    ch->step++;
  }
#undef QW
}
