/*
 * sx_oracle.h -- CPU oracle for the 6502 search core.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the Go CPU
 * algorithm of oisee/6502-optimizer (pkg/cpu, pkg/inst, pkg/search), used by
 * the test-suite, the smoke check and the bench's CPU-baseline leg as the
 * checker.  Nothing in the product library links or calls it.
 *
 * Each function cites the reference file:line it restates.  Pinning of each
 * function (known answers, closed forms, brute force, published counts) lives
 * in tests/test_oracle_*.py.
 */
#ifndef SX_ORACLE_H
#define SX_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Opcode ids: pkg/inst/instruction.go:42-122 (iota order). */
enum {
  SXO_TAX = 0, SXO_TAY, SXO_TXA, SXO_TYA, SXO_TSX, SXO_TXS,
  SXO_INX, SXO_INY, SXO_DEX, SXO_DEY,
  SXO_ASL_A, SXO_LSR_A, SXO_ROL_A, SXO_ROR_A,
  SXO_CLC, SXO_SEC, SXO_CLD, SXO_SED, SXO_CLI, SXO_SEI, SXO_CLV,
  SXO_PHA, SXO_PLA, SXO_PHP, SXO_PLP, SXO_NOP,
  SXO_LDA_IMM, SXO_LDX_IMM, SXO_LDY_IMM, SXO_ADC_IMM, SXO_SBC_IMM,
  SXO_AND_IMM, SXO_ORA_IMM, SXO_EOR_IMM, SXO_CMP_IMM, SXO_CPX_IMM, SXO_CPY_IMM,
  SXO_LDA_M, SXO_LDX_M, SXO_LDY_M, SXO_STA_M, SXO_STX_M, SXO_STY_M,
  SXO_ADC_M, SXO_SBC_M, SXO_AND_M, SXO_ORA_M, SXO_EOR_M,
  SXO_CMP_M, SXO_CPX_M, SXO_CPY_M, SXO_INC_M, SXO_DEC_M,
  SXO_ASL_M, SXO_LSR_M, SXO_ROL_M, SXO_ROR_M, SXO_BIT_M,
  SXO_OPCOUNT
};

/* State byte order A,X,Y,P,S,M,S0,S1: pkg/cpu/state.go:11-16. */
enum { SXO_A = 0, SXO_X, SXO_Y, SXO_P, SXO_S, SXO_M, SXO_S0, SXO_S1 };

/* Register masks: pkg/search/verifier.go:134-143. */
enum {
  SXO_RA = 1, SXO_RP = 2, SXO_RX = 4, SXO_RY = 8,
  SXO_RS = 16, SXO_RM = 32, SXO_RS0 = 64, SXO_RS1 = 128
};

/* Instruction word: op | imm << 16 (Go Instruction{Op u16, Imm u16}). */
#define SXO_WORD(op, imm) ((uint32_t)(op) | ((uint32_t)(imm) << 16))
#define SXO_OP(w) ((uint16_t)((w) & 0xFFFFu))
#define SXO_IMM(w) ((uint16_t)((w) >> 16))

/* Rule record, identical in layout to sx_rule in include/sx6502.h. */
typedef struct sxo_rule {
  uint32_t src[4];
  uint32_t rep[4];
  uint64_t target_index;
  uint8_t src_len, rep_len, dead_flags, pass;
  int16_t bytes_saved, cycles_saved;
} sxo_rule;

typedef struct sxo_stats {
  uint64_t checked;     /* candidates visited, counted like worker.go:126 */
  uint64_t qc_hits;     /* pairs that passed the byte filter, prune and QC */
  uint64_t mid_pass;    /* pairs that passed MidCheck (all-hits mode only) */
  uint64_t exh_calls;   /* ExhaustiveCheck invocations */
  uint64_t found;       /* rules emitted */
} sxo_stats;

/* ISA tables ------------------------------------------------------------ */
int sxo_has_imm(int op);
int sxo_byte_size(int op, int model);         /* model 0 = Go, 1 = CUDA tool */
int sxo_cycles(int op);
const char *sxo_mnemonic(int op);
int sxo_disasm(uint32_t w, char *buf, int cap);
int sxo_instruction_table(uint32_t *out, int order); /* 0 = Go, 1 = CUDA; returns 2863 */

/* Executor --------------------------------------------------------------- */
int sxo_exec(uint8_t *st, int op, uint16_t imm);  /* returns cycles, -1 unknown op */
void sxo_exec_seq(uint8_t *st, const uint32_t *seq, int n);
/* states[8*i..] executes words[i], for i < n (test helper for brute force). */
void sxo_exec_many(uint8_t *states, const uint32_t *words, int64_t n);
const uint8_t *sxo_qc_vectors(void);   /* 8 x 8 bytes */
const uint8_t *sxo_mid_vectors(void);  /* 32 x 8 bytes */

/* Verifier --------------------------------------------------------------- */
int sxo_states_equal_masked(const uint8_t *a, const uint8_t *b, uint8_t dead);
int sxo_quick_check(const uint32_t *t, int tn, const uint32_t *c, int cn, uint8_t dead);
int sxo_mid_check(const uint32_t *t, int tn, const uint32_t *c, int cn, uint8_t dead);
/* mode 0 = Go (32-value reduced sweep for >= 3 extra registers), 1 =
 * complete by brute force over the whole domain, 2 = complete by projection
 * over the dependencies of sxo_dep_units.  sxo_exhaustive_cex also writes a
 * disagreeing domain state to cex (when not NULL) if the pair fails. */
int sxo_exhaustive_check(const uint32_t *t, int tn, const uint32_t *c, int cn,
                         uint8_t dead, int mode);
int sxo_exhaustive_cex(const uint32_t *t, int tn, const uint32_t *c, int cn,
                       uint8_t dead, int mode, uint8_t cex[8]);
/* Mode 1 (complete brute force) on nthreads threads, executing 512 states
 * per opcode dispatch; cex as above.  sxo_exec_block runs one word on n
 * states (8 bytes each). */
int sxo_exhaustive_brute(const uint32_t *t, int tn, const uint32_t *c, int cn, uint8_t dead, int nthreads,
                         uint8_t cex[8]);
void sxo_exec_block(uint8_t *states, int n, uint32_t w);
/* d[u] = input units output unit u of opcode op is a function of (units: state
 * byte r = A 0, X 1, Y 2, S 4, M 5, S0 6, S1 7; flag bit b of P = 8 + b; an
 * unchanged unit depends on itself).  Returns -1 for an unknown opcode. */
int sxo_dep_units(int op, uint16_t d[16]);
uint8_t sxo_flag_diff(const uint32_t *t, int tn, const uint32_t *c, int cn);
void sxo_fingerprint(const uint32_t *seq, int n, uint8_t fp[64]);
int sxo_op_reads(int op);
int sxo_op_writes(int op);
int sxo_regs_read(const uint32_t *seq, int n);
int sxo_extra_regs(int reads, int out[6]);  /* setReg offsets 1..6 */

/* Pruner / enumerator ------------------------------------------------------ */
int sxo_should_prune(const uint32_t *seq, int n);
/* Pruned sequences of length len (1 or 2) in Go order.  Returns the count;
 * writes at most cap sequences (len words each) when out != NULL. */
int64_t sxo_enumerate_pruned(int len, uint32_t *out, int64_t cap);

/* Search driver ------------------------------------------------------------ */
/* hits: 0 = first hit in Go order, two passes when dead != 0 (worker.go);
 *       1 = all hits, single masked pass, QC -> Mid -> Exh (the CUDA tool). */
typedef struct sxo_search_cfg {
  int target_len;
  int max_cand_len;   /* 1, 2 or 3 (3: first hit only; length 3 is enumerated on the fly) */
  int bytes_model;    /* 0 Go, 1 CUDA */
  int hits;           /* 0 first, 1 all */
  int exh_mode;       /* as sxo_exhaustive_check: 0 Go reduced, 1 complete (brute force), 2 complete (projection) */
  int use_mid;        /* first-hit mode: 1 = run MidCheck as a pre-filter */
  int nthreads;
  uint8_t dead;
  int include_empty;  /* 1: also check the empty sequence per target (rules with rep_len 0, a separate stream) */
} sxo_search_cfg;

int sxo_search(const sxo_search_cfg *cfg, const uint32_t *targets, int64_t ntargets,
               sxo_rule *out, int64_t cap, int64_t *nout, sxo_stats *stats);

/* The pairs of the CUDA tool's full length-2 run (all pruned length-2
 * targets x the 2,863 single instructions, CUDA byte sizes, --dead-flags 0,
 * cuda/6502_search_v2.cu:848-977) that pass its byte filter, candidate
 * pruner, QuickCheck and MidCheck and read at least min_extra extra
 * registers (X, Y, M, S, S0, S1) between them -- with min_extra 3, the pairs
 * the tool sends to its CPU reduced sweep (cu:650,945-951).  out: (t0, t1, c)
 * words per pair in Go order of (target, candidate); returns the count. */
int64_t sxo_cuda_fallback_pairs(int min_extra, uint32_t *out, int64_t cap, int nthreads);

/* STOKE ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel Random Numbers: As
 * Easy as 1, 2, 3", SC'11). */
void sxo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Chain state of the STOKE search as defined for this build (the reference's
 * move set and verify trigger, pkg/stoke/mutator.go and search.go, with the
 * exact Metropolis-Hastings ratio, Philox randomness and the Hamming cost). */
typedef struct sxo_chain {
  uint32_t seq[10];
  int32_t len;
  int32_t cost10;     /* 10 x Hamming distance over the 32 Mid vectors + bytes */
  double temp;
  uint64_t step;
  uint64_t accepted;
  int32_t hits;       /* steps whose accepted state had Hamming 0 and fewer bytes */
  int32_t first_hit_step;
} sxo_chain;

/* cost10 of a candidate against a target (Go byte sizes). */
int sxo_stoke_cost10(const uint32_t *t, int tn, const uint32_t *c, int cn, uint8_t dead);
/* Initialise a chain at the target. */
void sxo_stoke_init(const uint32_t *t, int tn, double t0, uint8_t dead, sxo_chain *ch);
/* Run nsteps steps.  weights: NULL = uniform over the 2,863 instructions in
 * Go order, else 2,863 uint32 weights summing to 2^32 (the proposal
 * distribution of an instruction).  max_len = max(tn + 2, 10). */
void sxo_stoke_steps(const uint32_t *t, int tn, uint8_t dead, uint32_t chain_id, uint64_t seed,
                     double decay, const uint32_t *weights, int nsteps, sxo_chain *ch);

#ifdef __cplusplus
}
#endif
#endif
