/*
 * rqx.h -- C ABI of librqx.so, the gfx950 (MI355X) implementation of the
 * SQD subspace Pauli-sum action and its helpers.
 *
 * Conventions
 *   - States are uint64 bitstrings: bit q is qubit q.  The subspace index is
 *     the sorted, unique list of real states; rows [n_states, n_total) of every
 *     vector are filler slots (the reference fills them with 0xFF rows,
 *     rqutils/sqd.py:473-478) and always receive 0.
 *   - Pauli terms are grouped by X mask (groups sorted ascending by xmask,
 *     rqutils/paulis/symplectic.py:98).  Group j owns terms
 *     [grp_off[j], grp_off[j+1]) with Z masks zmask[] and phased coefficients
 *     coef[] = alpha * (-i)^(popcount(xmask & zmask) & 3)
 *     (symplectic.py:109-111), stored as interleaved (re, im) doubles.
 *   - The action is  y[i] = sum_j C_j(s_i) * x[rank(s_i ^ xmask_j)],
 *     C_j(s) = sum_k coef_k * (-1)^popcount(s & zmask_k)   (sqd.py:65-86).
 *   - Vector pointers are device pointers owned by the caller; every call is
 *     asynchronous on `stream` (a hipStream_t, NULL = default stream).
 *   - Return value: RQX_OK or a negative status; rqx_last_error() gives the
 *     message (thread local).  No C++ exception crosses this ABI.
 *   - A plan may be used by one host thread / stream at a time.
 */
#ifndef RQX_H
#define RQX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rqx_plan rqx_plan;
typedef void *rqx_stream_t;

enum {
    RQX_OK = 0,
    RQX_ERR_ARG = -1,     /* invalid argument (sizes, ordering, null pointer) */
    RQX_ERR_HIP = -2,     /* HIP runtime error */
    RQX_ERR_DTYPE = -3,   /* real vector dtype requested for a complex Hamiltonian */
    RQX_ERR_LIMIT = -4    /* size beyond a documented limit */
};

enum { RQX_F64 = 0, RQX_C128 = 1, RQX_F32 = 2, RQX_C64 = 3 };

typedef struct {
    int32_t n_qubits;
    int32_t split_bits;     /* L: states are segmented by s >> L, searched by s & (2^L-1) */
    int64_t n_states;       /* real (unique) states */
    int64_t n_total;        /* vector length incl. filler slots */
    int32_t n_groups;       /* J */
    int32_t n_classes;      /* distinct xmask >> L */
    int64_t n_terms;        /* T */
    int32_t n_segments;     /* distinct s >> L */
    int32_t n_tiles;        /* work tiles of <= 256 rows within one segment */
    int32_t max_segment;    /* largest segment (rows) */
    int32_t has_diagonal;   /* group 0 has xmask == 0 */
    int32_t is_real;        /* every coefficient is real */
    int32_t max_group_terms;
    int32_t n_problems;     /* 1, or the number of problems of a batched plan */
    int32_t persistent;     /* matvec launches persistent workgroups on a tile queue */
    int32_t blocks_per_cu;  /* persistent grid: workgroups per CU */
    int32_t n_cu;           /* compute units of the device */
} rqx_plan_info;

/* One problem of a (batched) plan: its rows, terms and groups inside the plan. */
typedef struct {
    int64_t row_offset;     /* first row of the problem in the plan's vectors */
    int64_t n_states;
    int64_t term_offset;    /* first term in the plan's term order (rqx_grad_coef) */
    int64_t n_terms;
    int32_t group_offset;
    int32_t n_groups;
    int32_t n_qubits;
    int32_t split_bits;
    int32_t n_segments;
    int32_t n_classes;
    int32_t n_tiles;
    int32_t max_segment;
    int32_t has_diagonal;
    int32_t is_real;
    double cost;            /* cost-model estimate of one matvec (arbitrary units) */
    int32_t section_bits;   /* hit-cache section split: groups whose X mask has no bit below it
                               form the first section (>= split_bits; larger when the states are a
                               product of low and high strings at that bit) */
    int32_t alpha_rows;     /* > 0: the states are the product of alpha_rows low strings (below
                               section_bits) and rows / alpha_rows high strings, and the matvec
                               takes the hits of the groups that flip only low bits from
                               per-low-string lists */
    int32_t list_mode;      /* 0: every class searched; 1: low-bit-only groups listed; 2: every
                               group listed (low-bit-only, high-bit-only and mixed groups through
                               per-string lists and a pattern map), no search at all */
    int32_t reserved;
} rqx_problem_info;

/* Library version (major*10000 + minor*100 + patch). */
int rqx_version(void);

/* Last error message of the calling thread ("" if none). */
const char *rqx_last_error(void);

/*
 * Build a plan: upload the grouped term tables and the sorted subspace index
 * to the current device, choose the segment split, precompute the x == 0
 * group's diagonal, the per-segment lists of classes whose target segment
 * exists, the sign tables of groups whose terms differ only on their X bits,
 * and (splits of at most 18 bits) per-segment hash tables of the low keys,
 * built on host threads.  All input pointers are HOST pointers.
 * split_bits < 0 selects the split with the host cost model.
 * Replaces the setup half of run_sqd (rqutils/sqd.py:341-351: uniquified
 * states + the J x N xsource cache) -- no J x N cache is built.
 * Limits: 1 <= n_qubits <= 64, n_total < 2^31 (sqd.py:140-145).
 */
int rqx_plan_create(rqx_plan **out, int32_t n_qubits,
                    const uint64_t *xmask, const int64_t *grp_off, int32_t n_groups,
                    const uint64_t *zmask, const double *coef, int64_t n_terms,
                    const uint64_t *sorted_states, int64_t n_states, int64_t n_total,
                    int32_t split_bits, rqx_stream_t stream);

/*
 * Batched plan over n_problems independent (Hamiltonian, subspace) pairs, each
 * given as in rqx_plan_create (arrays of HOST pointers, one entry per problem;
 * no filler rows: n_total = sum of n_states).  Every kernel of the plan then
 * covers all problems in one launch: vectors are the concatenation of the
 * problems' rows, term outputs the concatenation of their terms.  The matvec
 * runs persistent workgroups that take tiles from a queue, largest first.
 * No counterpart in the reference (one Hamiltonian per run_sqd call,
 * rqutils/sqd.py:317-323).
 */
int rqx_plan_create_batch(rqx_plan **out, int32_t n_problems, const int32_t *n_qubits,
                          const uint64_t *const *xmask, const int64_t *const *grp_off, const int32_t *n_groups,
                          const uint64_t *const *zmask, const double *const *coef, const int64_t *n_terms,
                          const uint64_t *const *sorted_states, const int64_t *n_states, int32_t split_bits,
                          rqx_stream_t stream);

/*
 * rqx_plan_create with the sorted unique states already in DEVICE memory (e.g.
 * the output of rqx_sort_unique_u64); the term tables are host pointers as in
 * rqx_plan_create.  The passes over the states -- validation, the segment
 * statistics of every split, the segment starts of the candidate splits, the
 * section split, the low keys and the per-segment hash tables -- run on the
 * device; only per-segment arrays are copied back for the cost model.  The
 * plan equals the one rqx_plan_create builds from the same states.
 * Replaces the setup half of run_sqd (rqutils/sqd.py:341-356) after
 * uniquify_states (sqd.py:341, 445-482), with no host round trip of the states.
 */
int rqx_plan_create_device(rqx_plan **out, int32_t n_qubits,
                           const uint64_t *xmask, const int64_t *grp_off, int32_t n_groups,
                           const uint64_t *zmask, const double *coef, int64_t n_terms,
                           const uint64_t *sorted_states, int64_t n_states, int64_t n_total,
                           int32_t split_bits, rqx_stream_t stream);

/*
 * Checksums (64-bit FNV-1a over 64-bit words) of the plan's index arrays, in
 * this order: keys, low keys, segment high parts, segment offsets, class high
 * parts, class group offsets, tile first rows, tile segments, tile problems,
 * problem descriptors, hash-table offsets, hash-table sizes, hash tables,
 * target-list offsets, target lists, FP64 diagonal.  out: RQX_DIGEST_ARRAYS
 * host words.  Two plans with equal digests hold the same index (e.g. one built
 * by rqx_plan_create and one by rqx_plan_create_device from the same states).
 * No counterpart in the reference; for reproducibility checks.
 */
enum { RQX_DIGEST_ARRAYS = 16 };
int rqx_plan_digest(const rqx_plan *p, uint64_t *out, rqx_stream_t stream);

int rqx_plan_destroy(rqx_plan *plan);

int rqx_plan_get_info(const rqx_plan *plan, rqx_plan_info *out);

/* Offsets and index statistics of problem k (0 <= k < n_problems). */
int rqx_plan_get_problem(const rqx_plan *plan, int32_t k, rqx_problem_info *out);

/*
 * Launch shape of the tile kernels: persistent (work queue) or one workgroup
 * per tile, and the persistent grid in workgroups per CU (0 = occupancy limit).
 * Used by the autotuner.
 */
int rqx_plan_set_launch(rqx_plan *plan, int32_t persistent, int32_t blocks_per_cu);

/*
 * idx[j * n_total + i] = rank of s_i ^ xmask_j among the real states, or -1
 * (also -1 for filler rows).  Replaces get_xsource (rqutils/sqd.py:485-548)
 * for all groups, i.e. the scan at sqd.py:347-351.  idx: device, J*n_total int32.
 * vals (optional, may be NULL): device, J*n_total complex interleaved doubles;
 * receives the matrix element C_j(s_i) where idx >= 0 and 0 elsewhere, i.e.
 * the (columns, elements) pair that hproj scans for (rqutils/sqd.py:300-307).
 */
int rqx_lookup(const rqx_plan *plan, int32_t *idx, double *vals, rqx_stream_t stream);

/*
 * diag[i] = C_0(s_i) for the x == 0 group (complex, interleaved doubles,
 * n_total entries, 0 on filler rows and when no such group exists).
 * Replaces get_diagonal(z[0], c[0], states_u) (rqutils/sqd.py:603-628,410).
 */
int rqx_diagonal(const rqx_plan *plan, double *diag, rqx_stream_t stream);

/*
 * y = H_S x for one vector of length n_total (device pointers, dtype one of
 * RQX_F64 / RQX_C128 / RQX_F32 / RQX_C64; F32/C64 accumulate in FP32).
 * Replaces apply_h_x_cached + make_x_cached_matvec on one device
 * (rqutils/sqd.py:694-709,746-751) and every other cache strategy
 * (sqd.py:643-743): all give the same numbers.  x and y must not alias.
 * Reads the hit cache of the dtype's precision when one is built (see
 * rqx_plan_build_cache), else searches every (row, group) pair on the fly;
 * both give the same values.
 */
int rqx_matvec(const rqx_plan *plan, const void *x, void *y, int32_t dtype, rqx_stream_t stream);

/* qubits of the largest full space rqx_full_apply serves */
enum { RQX_FULL_MAX_QUBITS = 40 };

/*
 * Full-space action on n_cols right-hand sides: y[b * n_cols + c] =
 * sum_j C_j(b) x[(b ^ xmask_j) * n_cols + c] for every basis state b of 2^n
 * (n = the plan's qubits; x, y: device arrays of 2^n x n_cols entries of dtype,
 * row major, must not alias).  Uses only the plan's term tables (the plan may
 * hold any subspace, or none); no index is searched.  Replaces
 * PauliSumXZ.matmul (rqutils/paulis/symplectic.py:130-149): the same gather
 * rhs[indices ^ xsig] times the group's signed coefficient sum at the output
 * index, for all columns in one launch.  Single-problem plans; at most
 * RQX_FULL_MAX_QUBITS qubits (the caller allocates 2^n x n_cols entries).
 */
int rqx_full_apply(const rqx_plan *plan, const void *x, void *y, int64_t n_cols, int32_t dtype, rqx_stream_t stream);

/*
 * counts[0] = structural nonzeros (row, group) hits incl. the diagonal,
 * counts[1] = (row, group) pairs resolved at the segment level (the fine probes),
 * counts[2] = (tile, target segment) pairs visited (each stages one target).
 * counts: device, 3 x int64 (overwritten).
 */
int rqx_count(const rqx_plan *plan, int64_t *counts, rqx_stream_t stream);

/*
 * Every nonzero of the projected Hamiltonian as (row, column, element)
 * triplets: the diagonal (x == 0 group) and every hit, elements complex
 * (interleaved doubles).  Replaces the J x N columns/elements scan of hproj
 * (rqutils/sqd.py:300-314) with output proportional to the nonzeros.  rows,
 * cols, vals: device arrays of `capacity` entries (rqx_count's counts[0] is
 * the exact number); *count (device) receives the number of nonzeros, also
 * when it exceeds capacity (then only `capacity` are written).  The order of
 * the triplets is unspecified (one atomic claim per wave).
 */
int rqx_plan_hits(const rqx_plan *plan, int64_t *rows, int32_t *cols, double *vals, int64_t capacity,
                  int64_t *count, rqx_stream_t stream);

/*
 * grad[k] = sum_i Re( conj(gy_i) * (P_k x)_i ) for every term k (plan term
 * order), i.e. the gradient of Re<gy, H x> with respect to the unphased real
 * coefficient alpha_k of term k, where (P_k x)_i = (-i)^{x.z_k}
 * (-1)^{popcount(s_i & z_k)} x[rank(s_i ^ x_k)].  x, gy: device vectors of the
 * given dtype; grad: device, n_terms doubles (overwritten).  No counterpart in
 * the reference (JAX autodiff cannot differentiate get_diagonal's while_loop,
 * rqutils/sqd.py:611-628, in reverse mode); it is the backward of
 * rqx_matvec with respect to the coefficients.
 */
int rqx_grad_coef(const rqx_plan *plan, const void *x, const void *gy, double *grad, int32_t dtype,
                  rqx_stream_t stream);

/*
 * out[k] = Re <psi| P_k |psi> (unnormalised) for every term: the Hellmann-Feynman
 * derivative dE/dalpha_k of the ground energy at a normalised ground vector psi.
 * Same layout as rqx_grad_coef (it is rqx_grad_coef(psi, psi)).
 */
int rqx_expect_terms(const rqx_plan *plan, const void *psi, double *out, int32_t dtype, rqx_stream_t stream);

/*
 * Replace the phased coefficients of every term (device pointer, n_terms complex
 * interleaved doubles, plan term order) and rebuild the x == 0 diagonal, so a
 * plan can follow a changing coefficient vector (autograd, parameter scans).
 * is_real must state whether every new coefficient is real.
 */
int rqx_plan_update_coef(rqx_plan *plan, const double *coef, int32_t is_real, rqx_stream_t stream);

/*
 * Hit cache: the reference's cache_strategy (1, 2) (rqutils/sqd.py:343-407: source
 * indices per X group from get_xsource, coefficient sums per group from
 * get_diagonal, applied by apply_h_xz_cached, sqd.py:731-743), compacted to the
 * hits only.  Stored as sliced ELL: 64 rows per slice, each slice in two
 * sections (hits of groups that flip no bit below the split, whose gathers
 * coalesce across a slice, and the rest), each padded to its widest row.  A
 * level of the first section whose 64 source rows are the slice's rows shifted
 * by one offset (product subspaces: almost all) stores that offset instead of
 * 64 columns, and one whose 64 values are equal by construction (one group
 * evaluated the same way for every row, rows agreeing on its Z bits) stores one
 * value.  Entries are written in the order the matrix-free kernel accumulates
 * them and with the same values, so rqx_matvec gives identical results with or
 * without the cache.
 *
 * rqx_plan_build_cache: precision 32 (used by RQX_F32 / RQX_C64 matvecs), 64
 * (RQX_F64 / RQX_C128), or 0 for both in one value pass over the hits with
 * shared columns and level codes; values are real when the Hamiltonian is.
 * Allocates (device memory; synchronous); RQX_ERR_LIMIT if it would need more
 * than max_bytes (max_bytes <= 0: 90% of the free device memory), counting the
 * scratch of the build (the columns of the padded layout, the groups of the
 * first section's entries).  Rebuilding replaces a cache of the same precision
 * (both, for caches built together).  rqx_plan_update_coef rewrites the values
 * on its stream in place (no allocation), or releases the cache if the
 * coefficients change between real and complex.
 */
int rqx_plan_build_cache(rqx_plan *plan, int32_t precision, int64_t max_bytes, rqx_stream_t stream);

/*
 * Release the cache of one precision (32 or 64), or both (0); caches built
 * together are released together.  Synchronises the device.  A matvec captured
 * in a HIP graph keeps the path chosen at capture (the cache's arrays when one
 * was built): release or rebuild a cache only when no such graph will replay.
 */
int rqx_plan_drop_cache(rqx_plan *plan, int32_t precision);

/*
 * out[0..14) = {built, stored hits (nnz), stored values, device bytes, slices,
 * complex values, stored columns, first-section levels, second-section values
 * (incl. padding), first-section values of levels without one value for the
 * whole slice, second-section hits (the other second-section entries are
 * padding), padding entries among those first-section values, padding entries
 * among the stored first-section columns, first-section levels of padding only}
 * for the cache of one precision (for caches built together: the 64-bit cache
 * counts the shared arrays, the 32-bit its values).
 */
int rqx_plan_cache_info(const rqx_plan *plan, int32_t precision, int64_t *out);

/*
 * Whether rqx_matvec uses a built cache: 0 never (the matrix-free kernels), 1
 * always, 2 (a new plan's default) the faster path: rqx_plan_build_cache times
 * the cached and the matrix-free matvec once for each complex dtype its cache
 * serves whose matrix-free matvec runs the product kernel (every problem fully
 * listed), and keeps the matrix-free kernel where it was faster (real dtypes keep
 * the cache).  Both paths give the same bits.  No reference counterpart: cache_strategy picks the cache in
 * rqutils/sqd.py:343-407 and this refines that choice.
 */
int rqx_plan_use_cache(rqx_plan *plan, int32_t enable);

/*
 * The kernel rqx_matvec runs for dtype now: *path = 2 the hit cache, 1 the
 * product kernel (matrix-free, every hit from the lists), 0 the matrix-free tile
 * kernels.  No reference counterpart (reporting only).
 */
int rqx_plan_matvec_path(const rqx_plan *plan, int32_t dtype, int32_t *path);

/* doubles of per-problem work space of rqx_locg_step, and the slot of it that the
 * step reads but never writes: the scale of the relative residual (below) */
enum { RQX_LOCG_WORK = 168, RQX_LOCG_SCALE = 164 };

/*
 * One LOCG iteration for every problem of the plan (a batched plan, or a single
 * plan as one problem), with no host synchronisation, so a caller can capture
 * it in a HIP graph.  The step is the reference's 3 x 3 Rayleigh-Ritz body
 * (rqutils/ground_locg.py:197-423: residual, direction orthogonal to x and the
 * previous direction, Rayleigh-Ritz on [x, y, p], x and y updated and
 * re-orthonormalised) with the images A x, A y carried instead of recomputed,
 * so it costs one matvec.  No counterpart in the reference (one problem per
 * call, four matvecs per iteration).
 *   dtype: the vectors' type, any of the four (FP32 vectors: the bulk phase
 *     of a mixed solve stored in low precision; the inner products, Rayleigh-
 *     Ritz and normalisations are FP64 either way); low != 0 (RQX_F64 /
 *     RQX_C128 vectors only): the matvec runs in RQX_F32 / RQX_C64 on low
 *     copies of p and A p;
 *   vec: device vectors {x, Ax, y, Ay, p, Ap, p_low, Ap_low} over all rows
 *     (p_low, Ap_low only with low); x, Ax, y, Ay are updated in place;
 *   theta, rel: device, one double per problem: <x, A x> and |A x - theta x| /
 *     max(|theta|, scale) at the start of the step; active: device, one byte per
 *     problem, cleared when rel <= tol; an inactive problem's x and A x do not change;
 *   work: device, RQX_LOCG_WORK doubles per problem; problem k's
 *     work[k * RQX_LOCG_WORK + RQX_LOCG_SCALE] is the caller's scale (e.g. |A x0|
 *     for the normalised start vector x0, so that a zero eigenvalue still
 *     converges; 0: |theta| alone), the rest is scratch.
 */
int rqx_locg_step(const rqx_plan *plan, int32_t dtype, int32_t low, void *const *vec, double *theta, double *rel,
                  uint8_t *active, double *work, double tol, rqx_stream_t stream);

/* problems of a plan that rqx_locg_step_static serves (its work space: 5.5 MB per GPU) */
enum { RQX_LOCG_STATIC_PROBLEMS = 4096 };

/*
 * rqx_locg_step without a work-space argument, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  The work space is
 * a file-scope device array of the library (RQX_LOCG_STATIC_PROBLEMS x
 * RQX_LOCG_WORK doubles, one per GPU, held until the process ends); the call
 * allocates nothing.  scale: device, one double per problem (the residual scale
 * of rqx_locg_step's work slot, e.g. |A x0|), or NULL for 0; it is copied into
 * the work space on every call.  Every call on one GPU shares that array and
 * leaves its contents behind: calls that overlap on one GPU (other streams,
 * other threads, other plans) race.  The plan's device must be the current
 * device.  RQX_ERR_LIMIT for plans of more than RQX_LOCG_STATIC_PROBLEMS
 * problems.  Otherwise as rqx_locg_step.
 */
int rqx_locg_step_static(const rqx_plan *plan, int32_t dtype, int32_t low, void *const *vec, double *theta,
                         double *rel, uint8_t *active, const double *scale, double tol, rqx_stream_t stream);

/*
 * Lowest eigenpair of a small Hermitian matrix (n = 2 or 3), row-major,
 * complex interleaved, device pointers.  out = [theta, 0, k0re, k0im, ...].
 * Replaces eigenpair_2x2 / eigenpair_3x3 (rqutils/ground_locg.py:467-514),
 * same formulas, with a fallback when the cross-product eigenvector vanishes.
 */
int rqx_eigenpair_small(const double *mat, int32_t n, double *out, rqx_stream_t stream);

/* The same routine on HOST pointers (runs on the calling CPU thread). */
int rqx_eigenpair_small_host(const double *mat, int32_t n, double *out);

/* count independent eigenproblems: mats count x n x n, out count x (2 + 2n) (device). */
int rqx_eigenpair_batch(const double *mats, int32_t count, int32_t n, double *out, rqx_stream_t stream);

/*
 * Per-problem inner products over a plan's vectors:
 * out[2 * (k * n_pairs + q) + {0, 1}] = sum over rows of problem k of conj(a_q) * b_q.
 * a, b: HOST arrays of n_pairs (<= 9) DEVICE vector pointers; out: device doubles.
 * The batched counterpart of the jnp.sum(conj(u) * v) reductions of
 * rqutils/ground_locg.py:259-262,445.
 */
int rqx_batch_dots(const rqx_plan *plan, int32_t n_pairs, const void *const *a, const void *const *b, double *out,
                   int32_t dtype, rqx_stream_t stream);

/*
 * out[i] = sum_q coef[k * n_in + q] * v_q[i] for the rows i of problem k
 * (coef: device, complex interleaved doubles; real vectors use the real part).
 * v: HOST array of n_in (<= 4) DEVICE vector pointers; out may be one of them.
 * The batched counterpart of the vector updates of rqutils/ground_locg.py:336-341.
 */
int rqx_batch_combine(const rqx_plan *plan, int32_t n_in, const void *const *v, const double *coef, void *out,
                      int32_t dtype, rqx_stream_t stream);

/*
 * Packed state rows -> uint64 keys on the device.  rows: n_rows x row_bytes
 * bytes in device memory, each row big endian with a pad bit first, then
 * qubits n_qubits-1 .. 0, zero fill (row_bytes = ceil((n_qubits + 1) / 8));
 * rows whose pad bit is set are filler rows (sqd.py:250-255, 473-478).  keys:
 * device array of n_rows entries; keys[0..*n_real) are the real rows' states
 * (bit q = qubit q) in input order; n_real: host pointer (the call
 * synchronises `stream`).  Replaces the host unpacking of run_sqd's packed
 * states (rqutils/sqd.py:255 builds them, sqd.py:341 uniquifies them).
 */
int rqx_pack_to_u64(const uint8_t *rows, int64_t n_rows, int32_t row_bytes, int32_t n_qubits,
                    uint64_t *keys, int64_t *n_real, rqx_stream_t stream);

/*
 * Sort and deduplicate n uint64 keys on the device: out[0..*n_unique) is the
 * sorted unique list.  keys/out: device arrays of n entries (may not alias);
 * n_unique: host pointer (the call synchronises `stream`).
 * Replaces uniquify_states (rqutils/sqd.py:445-482) on uint64 keys.
 */
int rqx_sort_unique_u64(const uint64_t *keys, int64_t n, uint64_t *out, int64_t *n_unique,
                        rqx_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* RQX_H */
