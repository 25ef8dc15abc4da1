/*
 * sx6502.h -- C ABI of lib6502core, the MI355X search core for the 6502
 * superoptimizer (oisee/6502-optimizer).
 *
 * The reference has no FFI of its own; its in-process seams are Go functions
 * (cited per entry point below).  This ABI exposes them in batched form so
 * that one call carries millions of (target, candidate) pairs:
 *
 *   - plain pointers and sizes, no C++ or torch types;
 *   - instruction words are uint32 op | imm << 16, which is the memory layout
 *     of Go's inst.Instruction{Op uint16, Imm uint16} on little-endian hosts
 *     (pkg/inst/instruction.go:7-10), so a cgo caller can pass &seq[0];
 *   - every entry returns an int status (SX_OK or a negative SX_E* code),
 *     never throws, and checks every HIP call;
 *   - inputs are caller-owned and read only during the call; device memory is
 *     owned by the handle; outputs go to caller-provided buffers.
 *
 * A handle is single-owner: calls on one handle must not overlap.  Calls on
 * different handles are independent.
 */
#ifndef SX6502_H
#define SX6502_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SX_ABI_VERSION 3

/* Status codes. */
#define SX_OK 0
#define SX_EINVAL (-1)     /* bad argument (null pointer, bad length, bad mode) */
#define SX_EOPCODE (-2)    /* opcode >= 58 in an input sequence (Go panics: pkg/cpu/exec.go:185-186) */
#define SX_ECAPACITY (-3)  /* output buffer too small; the required count is written back */
#define SX_EHIP (-4)       /* a HIP runtime call failed; sx_last_error has the text */
#define SX_ENOMEM (-5)     /* device or host allocation failed */
#define SX_ENODEV (-6)     /* no usable gfx950 device */

/* Instruction word helpers. */
#define SX_WORD(op, imm) ((uint32_t)(op) | ((uint32_t)(imm) << 16))

typedef struct sx_handle sx_handle;

/* One rule.  Mirrors result.Rule (pkg/result/table.go:11-17) with fixed-size
 * sequences.  pass = 1 for an exact rule, 2 for a flag-relaxed rule (the
 * second pass of worker.go:180-245); dead_flags is FlagDiff for pass 2 in the
 * Go semantics and the command-line mask in the CUDA-tool semantics. */
typedef struct sx_rule {
  uint32_t src[4];
  uint32_t rep[4];
  uint64_t target_index; /* index into the target list of the call */
  uint8_t src_len, rep_len, dead_flags, pass;
  int16_t bytes_saved, cycles_saved;
} sx_rule;

/* Counters and stage times of the last enumeration (cf. worker.go:19-21 and
 * the CUDA tool's totals, cuda/6502_search_v2.cu:1045-1058). */
typedef struct sx_stats {
  uint64_t targets;      /* targets processed */
  uint64_t qc_pairs;     /* (target, candidate) pairs evaluated by QuickCheck */
  uint64_t qc_raw_hits;  /* pairs passing the QuickCheck predicate (any pass) */
  uint64_t qc_hits;      /* ... and the byte filter and candidate pruner */
  uint64_t mid_pass;     /* ... and MidCheck */
  uint64_t exh_pairs;    /* ExhaustiveCheck jobs run */
  uint64_t exh_states;   /* states evaluated by ExhaustiveCheck (counted on the device) */
  uint64_t exh_rounds;   /* first-hit rounds */
  uint64_t rules;        /* rules emitted */
  uint64_t overflow_retries; /* compaction-buffer overflows recovered by re-running a tile */
  double ms_setup, ms_qc, ms_mid, ms_exh, ms_total; /* device times (HIP events) and wall */
  double ms_qc_kernel, ms_exh_kernel; /* HIP events around the QuickCheck / exhaustive-sweep launches only
                                         (QuickCheck's: with MidCheck when the two run as one join) */
  uint64_t qc_launches, exh_launches;
  uint32_t qc_tile;          /* QuickCheck targets per work item (autotuned unless tile_targets is set) */
  uint32_t qc_blocks_per_cu; /* 0 = grid launch; else persistent launch with this many blocks per CU */
  uint64_t empty_rules;      /* rules of the empty-candidate stream (include_empty) */
  /* the exhaustive sweep's one-job-per-lane kernel (jobs of length-2 targets
     against single instructions with small plans); ms_exh_kernel,
     exh_launches cover the state-parallel kernel, exh_states both */
  double ms_exh_group_kernel;
  uint64_t exh_group_launches;
  uint64_t exh_group_jobs;   /* jobs it swept */
  uint64_t exh_group_states; /* states it evaluated for jobs still open */
} sx_stats;

/* Enumeration semantics. */
#define SX_BYTES_GO 0     /* pkg/inst/catalog.go:50-56: M ops are 1 byte */
#define SX_BYTES_CUDA 1   /* cuda/6502_search_v2.cu:437-440: M ops are 2 bytes */
#define SX_HITS_FIRST 0   /* first hit in Go order, exact pass then flag-relaxed pass (worker.go:119-245) */
#define SX_HITS_ALL 1     /* every hit of one masked pass (cuda/6502_search_v2.cu:865-974) */
#define SX_EXH_PROJECTED 0 /* complete sweep, split per output byte over its data dependencies */
#define SX_EXH_FULL 1      /* complete sweep over all 2^(9+8k) states of the domain */

typedef struct sx_enum_cfg {
  uint32_t struct_size;   /* sizeof(sx_enum_cfg) */
  int32_t target_len;     /* 2 or 3 */
  int32_t max_cand_len;   /* 1 or 2 (search.go:26-28: --max-target L gives L - 1 for every target length) */
  int32_t bytes_model;    /* SX_BYTES_* */
  int32_t hits;           /* SX_HITS_* */
  int32_t exh_mode;       /* SX_EXH_* */
  int32_t use_mid;        /* 1: MidCheck filter before ExhaustiveCheck */
  uint32_t dead_flags;    /* --dead-flags mask (0..255) */
  const uint32_t *targets;/* n_targets * target_len words, or NULL = all pruned
                             length-2 targets in Go order (target_len 2 only) */
  uint64_t n_targets;
  uint64_t first_target;  /* with targets == NULL: half-open range of the */
  uint64_t end_target;    /* pruned target index space (0, 0 = everything) */
  uint32_t tile_targets;  /* QuickCheck targets per tile, all-pairs grid launch, rounded up to a multiple
                             of 32 (0 = with use_mid, QuickCheck and MidCheck as one hash join; without,
                             the autotuned shape) */
  uint32_t survivor_cap;  /* survivor buffer entries per segment: QuickCheck's, split over 8 regions, or
                             with the join MidCheck's (0 = default); a full buffer re-runs the segment
                             with a larger one */
  int32_t include_empty;  /* 1: also check every target against the empty sequence (0 bytes, never
                             pruned).  Go never tries it (worker.go:123 starts at length 1), so its
                             rules (rep_len 0) form a stream of their own, fetched with
                             sx_fetch_empty_rules, and do not change the other rules; QuickCheck
                             pairs then count the empty slot too */
  uint32_t reserved;      /* 0 */
} sx_enum_cfg;
/* The size of sx_enum_cfg before include_empty; a caller passing it gets include_empty = 0. */
#define SX_ENUM_CFG_SIZE_V1 72u

/* Open a handle on HIP device `device`.  Replaces the process start of the
 * CUDA tool (cuda/6502_search_v2.cu:763-797) and the WorkerPool construction
 * (pkg/search/worker.go:25-35). */
int sx_open(int device, sx_handle **out);
int sx_close(sx_handle *h);

/* Copy the text of the last error on this handle into buf. */
int sx_last_error(sx_handle *h, char *buf, size_t cap);
int sx_abi_version(void);

/* Run one enumeration.  Replaces search.Run (pkg/search/search.go:22-54)
 * and, per target, WorkerPool.processTask / processTaskMasked
 * (pkg/search/worker.go:119-245); with SX_HITS_ALL and SX_BYTES_CUDA it
 * reproduces the CUDA tool's pipeline (cuda/6502_search_v2.cu:848-977).
 *
 * The rules are kept in the handle, sorted by (target_index, pass).  If
 * `out` is non-NULL and `cap` is large enough they are copied there; if it is
 * too small SX_ECAPACITY is returned and *n_out holds the required count;
 * sx_fetch_rules then copies them without re-running. */
int sx_enumerate(sx_handle *h, const sx_enum_cfg *cfg, sx_rule *out, uint64_t cap,
                 uint64_t *n_out, sx_stats *stats);
int sx_fetch_rules(sx_handle *h, sx_rule *out, uint64_t offset, uint64_t cap, uint64_t *n_copied);
/* The empty-candidate stream of the last enumeration (include_empty), sorted
 * by (target_index, pass): per target at most one rule, pass 1 when the
 * target equals the empty sequence under the exact checks, pass 2 when only
 * the flag-relaxed ones hold (dead_flags = FlagDiff); in the CUDA-tool
 * semantics one masked pass.  Same calling pattern as sx_fetch_rules. */
int sx_fetch_empty_rules(sx_handle *h, sx_rule *out, uint64_t offset, uint64_t cap, uint64_t *n_copied);

/* Number of pruned length-2 targets in Go order (7,633,612) and the words of
 * the range [first, first + n) of them (search.go:56-75 collectTasks). */
int sx_pruned_targets(sx_handle *h, uint64_t first, uint64_t n, uint32_t *out_words,
                      uint64_t *total);

/* Check n (target, candidate) pairs in bulk.  Sequences are up to 4 words,
 * stored 4 words per pair with explicit lengths.  stage_mask bit 0 =
 * QuickCheck, bit 1 = MidCheck, bit 2 = ExhaustiveCheck (complete).  Each
 * requested stage is evaluated on its own; out[i] gets bit s set when stage s
 * holds for pair i under the dead-flag mask.  Replaces
 * search.QuickCheck/MidCheck/ExhaustiveCheck(Masked) (pkg/search/
 * verifier.go:84-105,262,427-466) and the per-line check of verify-jsonl
 * (cmd/6502opt/main.go:434-439). */
#define SX_STAGE_QC 1u
#define SX_STAGE_MID 2u
#define SX_STAGE_EXH 4u
int sx_check_pairs(sx_handle *h, const uint32_t *t_words, const uint8_t *t_len,
                   const uint32_t *c_words, const uint8_t *c_len, uint64_t n,
                   uint32_t stage_mask, uint32_t dead_flags, int32_t exh_mode, uint8_t *out);

/* sx_check_pairs without any allocation.  No counterpart in the reference:
 * this variant is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It calls
 * no hipMalloc, hipFree or other allocating routine; its device buffers live
 * in one fixed array of about 10.4 MiB that each GPU the library runs on
 * holds until the process ends.  Same arguments and result bits as
 * sx_check_pairs; at most SX_NOALLOC_MAX_PAIRS pairs per call (more give
 * SX_EINVAL).  Every call on one GPU shares that GPU's array, so calls that
 * overlap on one GPU (from any handle or thread) race. */
#define SX_NOALLOC_MAX_PAIRS 65536u
int sx_check_pairs_noalloc(sx_handle *h, const uint32_t *t_words, const uint8_t *t_len,
                           const uint32_t *c_words, const uint8_t *c_len, uint64_t n,
                           uint32_t stage_mask, uint32_t dead_flags, int32_t exh_mode, uint8_t *out);

/* Shortest replacement of one target of 1..8 instructions, candidates of
 * length 1..max_cand_len (<= 4) enumerated on the device in Go order.
 * Replaces search.SearchSingle / SearchSingleMasked (pkg/search/search.go:
 * 76-106) behind `6502opt target` (cmd/6502opt/main.go:88-132), i.e. the
 * per-target loops of worker.go:119-245: out[0] is the exact pass (first
 * candidate in Go order, shortest length first, that passes the byte filter,
 * the pruner, QuickCheck and the complete ExhaustiveCheck), out[1] the
 * masked pass (dead_flags != 0: first candidate failing the exact QuickCheck,
 * passing the masked QuickCheck and ExhaustiveCheck, FlagDiff != 0; its
 * dead_flags field is that FlagDiff).  found = 0 where a pass has no rule.
 * scanned: candidates enumerated (worker.go:126 counts the same). */
typedef struct sx_target_result {
  uint32_t rep[4];
  uint8_t rep_len, found, pass, dead_flags;
  int16_t bytes_saved, cycles_saved;
} sx_target_result;
int sx_target_search(sx_handle *h, const uint32_t *target, int target_len, int max_cand_len, uint32_t dead_flags,
                     int bytes_model, sx_target_result out[2], uint64_t *scanned);

/* FlagDiff for n pairs (pkg/search/verifier.go:591-604), same layout. */
int sx_flag_diff(sx_handle *h, const uint32_t *t_words, const uint8_t *t_len,
                 const uint32_t *c_words, const uint8_t *c_len, uint64_t n, uint8_t *out);

/* Run n sequences (up to 4 words each) on n states (8 bytes each, A X Y P S
 * M S0 S1) on the device; states are updated in place.  Exposes cpu.Exec
 * (pkg/cpu/exec.go:8) for testing the device interpreter. */
int sx_exec_states(sx_handle *h, uint8_t *states, const uint32_t *words, const uint8_t *len,
                   uint64_t n);

/* The sweep plan of one pair: the number of sweeps, each sweep's varying
 * register set (bit r = state byte r) and the total number of states, for
 * exh_mode.  Host-side; used to audit the projection. */
int sx_exh_plan(const uint32_t *t, int tn, const uint32_t *c, int cn, uint32_t dead_flags,
                int32_t exh_mode, uint8_t sets[8], int32_t *n_sets, uint64_t *states);

/* STOKE search (pkg/stoke): chains_per_target Metropolis-Hastings chains per
 * target, each starting at its target.  Cost = 10 x Hamming distance of the
 * output state over the 32 Mid vectors + byte size (Go sizes), in tenths;
 * moves and weights of pkg/stoke/mutator.go:30-106 (replace 102, swap 52,
 * delete 51, insert 25, immediate 26, out of 256); instructions drawn from
 * integer weights summing to 2^32 (uniform when weights == NULL, else one
 * table of 2,863 weights per target, Go instruction order); acceptance with
 * the exact proposal ratio at temperature t0 * decay^step; randomness from
 * Philox4x32-10 with counter (step, chain, stream) and key seed.  A chain whose
 * accepted state matches the target on all 32 vectors with fewer bytes posts
 * it; posted candidates are verified with the complete ExhaustiveCheck (the
 * verify trigger of pkg/stoke/search.go:101-112).  Replaces stoke.Run
 * (pkg/stoke/search.go:32-176). */
typedef struct sx_stoke_cfg {
  uint32_t struct_size;      /* sizeof(sx_stoke_cfg) */
  uint32_t chains_per_target;
  uint64_t steps;            /* steps per chain */
  uint32_t steps_per_launch; /* 0 = 256 steps, doubling per launch up to 4,096 */
  int32_t stop_when_verified;/* stop a target's chains at its first verified rewrite */
  double t0, decay;
  uint64_t seed;
  uint32_t dead_flags;
  uint32_t pad;
  const uint32_t *weights;   /* NULL or n x 2,863 uint32, each row summing to 2^32 */
} sx_stoke_cfg;

typedef struct sx_stoke_result {
  uint32_t rep[10];
  uint8_t rep_len, verified;
  uint8_t dead_flags;        /* FlagDiff(target, rep) when dead_flags != 0, else 0
                              * (pkg/stoke/search.go:104-110) */
  uint8_t pad;
  int16_t bytes_saved, cycles_saved;
  uint32_t chain;            /* chain that found it */
  uint64_t step;             /* its step index */
  double seconds;            /* wall time from the start of the search to its verification */
} sx_stoke_result;

typedef struct sx_stoke_chain {
  uint32_t seq[10];
  int32_t len, cost10;
  uint64_t step, accepted;
} sx_stoke_chain;

/* targets: n x 8 words (lengths in lens, 1..8); out: n results; chains_out:
 * optional n * chains_per_target final chain states. */
int sx_stoke_run(sx_handle *h, const sx_stoke_cfg *cfg, const uint32_t *targets, const uint8_t *lens,
                 uint64_t n, sx_stoke_result *out, sx_stoke_chain *chains_out);

/* One verified rewrite of a collecting run. */
typedef struct sx_stoke_found {
  uint32_t target;           /* index into the targets */
  uint32_t chain;            /* chain that found it */
  uint64_t step;             /* its step index */
  double seconds;            /* wall time from the start of the search to its verification */
  uint32_t rep[10];
  uint8_t rep_len;
  uint8_t dead_flags;        /* FlagDiff(target, rep) when dead_flags != 0, else 0 */
  int16_t bytes_saved, cycles_saved;
  uint16_t pad;
} sx_stoke_found;

/* stoke.Run followed by stoke.Deduplicate (pkg/stoke/search.go:32-176,
 * 179-190): every chain runs cfg->steps steps; every verified rewrite is kept
 * once per (target, replacement), in the order found (launch, then target,
 * step, chain), and the chain that found it restarts from its target at t0
 * (search.go:150-153; here at the next launch boundary, where the host
 * verifies posts).  stop_when_verified is ignored.  The first min(cap,
 * *n_found) rewrites are copied to found; *n_found is the number found. */
int sx_stoke_collect(sx_handle *h, const sx_stoke_cfg *cfg, const uint32_t *targets, const uint8_t *lens,
                     uint64_t n, sx_stoke_found *found, uint64_t cap, uint64_t *n_found,
                     sx_stoke_chain *chains_out);

/* The per-opcode data-dependency table behind SX_EXH_PROJECTED, at flag
 * granularity: inputs and outputs are numbered A=0, X=1, Y=2, S=4, M=5, S0=6,
 * S1=7 and 8 + b for bit b of P; out[op * 16 + o] = bit set of the inputs that
 * output o of opcode op is a function of (0xFFFF = output o is not written).
 * Host-side. */
int sx_dep_table(uint16_t out[58 * 16]);

/* ---- learned STOKE proposal (no counterpart in the reference: the
 * reference's mutator draws instructions uniformly, pkg/stoke/mutator.go:108-121).
 *
 * A two-layer MLP maps a target sequence (context: 8 tokens, token = the Go
 * instruction index 0..2862 of search.AllInstructions order, 2863 = padding)
 * to a distribution over 2,871 outputs: the 2,863 instructions, then 8
 * pointers, output 2863 + j meaning "the instruction at context position j":
 *   X = [E[t0] .. E[t7]] (8 x 64 = 512), H = relu(X W1 + b1) (256),
 *   p = softmax(H W2 + b2) (2,871).
 * Parameters (fp32 master copy, row-major, in this order): E[2864][64],
 * W1[512][256], b1[256], W2[256][2871], b2[2871] (1,052,471 floats).
 * Products run on MFMA with bf16 operands (E, W1, W2, H and the
 * back-propagated dZ, dH rounded to bf16, round-to-nearest-even) and fp32
 * accumulation; biases, softmax and Adam in fp32.  Training minimises the
 * mean cross entropy of the label (an output index: an instruction of a
 * known rewrite of the context, or the pointer to one of its positions). */
typedef struct sx_mlp sx_mlp;

uint64_t sx_mlp_param_count(void);
/* initial weights from seed (xorshift64): E ~ U(-0.1, 0.1), W1 ~ U(+-1/sqrt(512)),
 * W2 ~ U(+-1/sqrt(256)), zero biases; Adam state cleared. */
int sx_mlp_create(sx_handle *h, uint64_t seed, sx_mlp **out);
int sx_mlp_destroy(sx_mlp *m);
int sx_mlp_get_params(sx_mlp *m, float *out);       /* sx_mlp_param_count() floats */
int sx_mlp_set_params(sx_mlp *m, const float *in);  /* also clears the Adam state */
/* Mean loss over the n examples and its gradient (n x 8 ctx, n labels). */
int sx_mlp_loss_grad(sx_mlp *m, const uint16_t *ctx, const uint16_t *labels, uint32_t n, float *loss,
                     float *grad);
/* `steps` Adam steps (beta 0.9/0.999, eps 1e-8, bias-corrected) on batches of
 * `batch` (<= 4096) examples drawn uniformly with replacement from the n
 * examples (Philox4x32-10, counter = (Adam step, row), key = seed).  The step
 * is one captured hipGraph replayed `steps` times.  losses: optional, one
 * batch loss per step. */
int sx_mlp_train(sx_mlp *m, const uint16_t *ctx, const uint16_t *labels, uint32_t n, uint32_t steps,
                 uint32_t batch, float lr, uint64_t seed, float *losses);
/* probs: n x 2,871 softmax outputs. */
int sx_mlp_predict(sx_mlp *m, const uint16_t *ctx, uint32_t n, float *probs);
/* Proposal tables of n targets for sx_stoke_cfg.weights: predict, add each
 * pointer's probability to the instruction at its position (in position
 * order, fp32; padding positions add nothing), then sx_proposal_weights. */
int sx_mlp_proposal(sx_mlp *m, const uint16_t *ctx, uint32_t n, double eps, uint32_t *out);
/* Proposal tables for sx_stoke_cfg.weights: row i = (1 - eps) * probs[i] +
 * eps / 2863 as uint32 weights summing to 2^32, every weight >= 1 (the
 * uniform fallback keeps every instruction reachable, so the chain stays
 * irreducible and every proposal ratio finite).  Host-side. */
int sx_proposal_weights(const float *probs, uint64_t n, double eps, uint32_t *out);
/* C = A B through the MLP's MFMA tile kernel (A: M x K, B: K x N bf16
 * row-major, C: M x N fp32); a check of the kernel's operand layout. */
int sx_mfma_gemm_test(sx_handle *h, const uint16_t *A, const uint16_t *B, int M, int N, int K, float *C);
/* Measurement helper (no counterpart in the reference): the device's int32
 * VALU rate in lane-operations per second, from a kernel of independent
 * v_xor_b32 chains on every CU (best of three runs, about 30 ms).  The bench
 * reports it beside the roofline peak. */
int sx_measure_int32_rate(sx_handle *h, double *ops_per_s);

#ifdef __cplusplus
}
#endif
#endif
