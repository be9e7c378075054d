/*
 * pqi.h -- C ABI of libpqi.so, the MI355X (gfx950) cut-evaluation and
 * bit-flip local-search library.
 *
 * The ABI replaces PyQrackIsing's OpenCL seam (pyqrackising/maxcut_tfim_util.py:9-301:
 * OpenCLContext, setup_opencl, make_*_buf) and the host wrappers that drive the
 * kernels in pyqrackising/kernels.cl.  Each entry point cites the reference
 * function whose work it takes over.
 *
 * Conventions (all from the reference):
 *   - bits are x in {0,1}^n, one uint8 per node for solver state; candidate
 *     batches are packed B x ceil(n/32) uint32 words, LSB-first
 *     (pyqrackising/maxcut_tfim_util.py:304-316);
 *   - cut(x) = sum_{i<j} w_ij [x_i != x_j]; energy E(x) = 2 cut(x) - sum_{i<j} w_ij
 *     (pyqrackising/maxcut_tfim_util.py:319-331);
 *   - selection is the first maximum (np.argmax);
 *   - dense W must be symmetric with a zero diagonal; CSR input must be strictly
 *     upper triangular (pyqrackising/maxcut_tfim_util.py:533-542).
 *
 * Every function returns PQI_OK (0) or a negative code; pqi_last_error() returns
 * a thread-local message.  No C++ exception crosses this boundary.  The caller
 * owns every host array; nothing is retained after a call returns.  A context
 * owns one HIP stream and is not re-entrant.
 */
#ifndef PQI_H
#define PQI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define PQI_API __attribute__((visibility("default")))
#else
#define PQI_API
#endif

#define PQI_OK 0
#define PQI_ERR_INVALID (-1)     /* bad argument (shape, dtype, range, symmetry) */
#define PQI_ERR_NO_DEVICE (-2)   /* no HIP device, or the device index is out of range */
#define PQI_ERR_HIP (-3)         /* a HIP runtime call failed */
#define PQI_ERR_OOM (-4)         /* device allocation failed */
#define PQI_ERR_UNSUPPORTED (-5) /* valid request outside what this build implements */

/* element types accepted for weights */
#define PQI_DTYPE_I8 1
#define PQI_DTYPE_I32 2
#define PQI_DTYPE_F32 3
#define PQI_DTYPE_F64 4

/* how a graph's weights are held on the device */
#define PQI_WEIGHTS_I8 1  /* integers in [-127, 127]: int8 storage, int32 fields, exact */
#define PQI_WEIGHTS_F32 3 /* fp32 storage, fp64 fields and re-scores */
#define PQI_WEIGHTS_F64 4 /* fp64 storage, fp64 fields and re-scores */

/* double-flip semantics */
#define PQI_DOUBLE_REFERENCE 0  /* pyqrackising/spin_glass_solver.py:55-92: thread t keeps its last pair */
#define PQI_DOUBLE_EXHAUSTIVE 1 /* pyqrackising/kernels.cl:186-231: best over all n(n-1)/2 pairs */

/* candidate-evaluation engine for pqi_eval_batch_ex */
#define PQI_ENGINE_AUTO 0  /* int8: MFMA (exact); fp32: split BF16 + FP64 re-score of the top 16; fp64: FP64 */
#define PQI_ENGINE_EXACT 1 /* int8: MFMA (exact); real weights: FP64 tiles */
#define PQI_ENGINE_BF16 2  /* fp32 weights only: split hi/lo BF16 MFMA, FP32 accumulation */

/* candidate-evaluation engine for Gray seeds */
#define PQI_SEEDS_AUTO 0
#define PQI_SEEDS_MFMA 1     /* E = rowsum((S W) .* S) on int8 MFMA (int8 weights only) */
#define PQI_SEEDS_LOWRANK 2  /* cut difference from the maintained field, O(|F|^2) per seed */

typedef struct pqi_ctx pqi_ctx;
typedef struct pqi_graph pqi_graph;
typedef struct pqi_batch pqi_batch;
typedef struct pqi_mpnn pqi_mpnn;

typedef struct {
    int64_t n;
    int32_t weights;  /* PQI_WEIGHTS_* */
    int32_t has_dense;
    int64_t nnz;      /* nonzeros of the full symmetric matrix */
    double w_up;      /* sum_{i<j} w_ij */
    int64_t device_bytes;
} pqi_graph_info;

/* Parameters of the spin_glass_solver local search
 * (pyqrackising/spin_glass_solver.py:424-439, 556-560, 603). */
typedef struct {
    int32_t is_spin_glass;
    int32_t double_mode;     /* PQI_DOUBLE_* */
    int64_t gray_iterations; /* default n*n */
    int64_t walkers;         /* thread_count = gray_seed_multiple in the reference */
    int64_t seed_multiple;   /* gray_seed_multiple; block_size = walkers * seed_multiple */
    int64_t max_rounds;      /* 0 = run until no phase improves */
    int32_t seed_engine;     /* PQI_SEEDS_* */
    int32_t flags;           /* PQI_FLAG_* */
} pqi_search_params;

/* pqi_search_params.flags */
#define PQI_FLAG_TIME_KERNELS 1 /* time the descent and main walk kernels with HIP events on the context stream
                                   (batched and learned searches: the descent launches of the whole call, the
                                   same totals in every instance record).  The walk is timed only where it is
                                   launched on its own (host-decided loop, segment graphs); on the int8 fast
                                   path and with the whole iteration captured as one graph it runs inside the
                                   iteration (events inside a graph cannot be timed on this ROCm), so
                                   walk_kernel_seconds and walk_launches stay 0 there */
#define PQI_FLAG_HOST_LOOP 2    /* decide each phase on the host (one synchronisation per phase) instead of
                                   on the device (one per loop iteration); same result */
#define PQI_FLAG_TEST_DRIFT 4   /* test hook (real weights): perturb the maintained field by
                                   1e-3 ((v mod 7) - 3) before the first round, which the round's exact
                                   recompute must detect and repair */
#define PQI_FLAG_FALLBACK 8     /* pqi_learned_search with targets and passes > 0: an instance that ends
                                   below its target is searched again by the plain batched loop from its
                                   own start, to the same target, within the same call; its record is then
                                   the plain search's (fallback = 1), with the learned phase's counters kept.
                                   The fallback runs the loop with max_rounds 0 when the caller's is < 0 */

typedef struct {
    double max_energy; /* reference bookkeeping: start value plus accepted gains */
    double cut_value;  /* exact re-score of the final state */
    double min_energy; /* -E (spin glass) or E (maxcut), exact re-score */
    int64_t rounds;
    int64_t single_accepts, double_accepts, seed_accepts, walk_accepts, reheat_accepts;
    int64_t single_evals, double_evals, seed_evals;
    int64_t gray_evals;          /* Gray steps of the reference walk, each resolved exactly */
    int64_t gray_evals_computed; /* of those, gains computed (the rest repeat a rejected step) */
    double seconds_single, seconds_double, seconds_seeds, seconds_walk, seconds_reheat;
    /* with PQI_FLAG_TIME_KERNELS: summed event-timed kernel durations and launch counts */
    double walk_kernel_seconds, descent_kernel_seconds;
    int64_t walk_launches, descent_launches;
    int64_t descent_evals; /* single-flip evaluations of the descents (single_evals less best-only checks) */
    /* learned phase (pqi_learned_search) */
    int64_t learned_accepts, learned_passes;
    double seconds_learned;
    /* real weights: the fp64 field maintained by flips is recomputed exactly from W at the start of
     * every round and replaced; entries that had moved by more than 1e-9 max(1, |h|) are counted */
    int64_t field_checks, field_repairs;
    double max_field_drift;
    /* learned phase: the smallest cut gain (unscaled) of an accepted learned move; +inf when none */
    double learned_min_gain;
    /* batched and learned searches: (instance, descent launch) pairs of the whole call, and those of them
     * whose descent moved (and wrote its state back); the same totals in every instance record */
    int64_t descent_instance_launches, descent_instance_writes;
    /* pqi_learned_search with PQI_FLAG_FALLBACK: 1 when the instance's result comes from the plain
     * fallback search, else 0 (descent timing totals and pairs then cover both searches of the call) */
    int64_t fallback;
    /* with PQI_FLAG_FALLBACK, a fallback record: the single-flip accepts of the learned search it replaces */
    int64_t replaced_single_accepts;
} pqi_search_stats;

/* Thread-local message for the last failing call on this thread. */
PQI_API const char* pqi_last_error(void);
PQI_API int pqi_version(void);

/* Number of visible HIP devices (0 with PQI_ERR_NO_DEVICE when none). */
PQI_API int pqi_device_count(int* count);

/* Replaces the import-time device pick, context and queue creation
 * (pyqrackising/maxcut_tfim_util.py:96-137). */
PQI_API int pqi_create(int device, pqi_ctx** ctx);
PQI_API int pqi_destroy(pqi_ctx* ctx);

/* Replaces make_G_m_buf (pyqrackising/maxcut_tfim_util.py:224-238).  W is n x n
 * row-major of the given dtype, symmetric, zero diagonal.  Integer-valued
 * weights in [-127,127] are held as int8 (exact); others as fp32/fp64. */
PQI_API int pqi_graph_dense(pqi_ctx* ctx, const void* W, int dtype, int64_t n, pqi_graph** g);

/* Replaces make_G_m_csr_buf (pyqrackising/maxcut_tfim_util.py:241-259).  The CSR
 * is strictly upper triangular (the reference's format); it is symmetrised on
 * the device side of the library. */
PQI_API int pqi_graph_csr(pqi_ctx* ctx, const int64_t* indptr, const int64_t* indices, const void* data,
                  int dtype, int64_t n, int64_t nnz, pqi_graph** g);
PQI_API int pqi_graph_free(pqi_graph* g);
PQI_API int pqi_graph_get_info(const pqi_graph* g, pqi_graph_info* info);

/* Candidate scoring.  Replaces run_cut_opencl + the calculate_cut(_sparse)
 * kernels (pyqrackising/maxcut_tfim.py:212-285, maxcut_tfim_sparse.py:227-306,
 * kernels.cl:44-94, 233-285).  bits: B x ceil(n/32) uint32.  values (optional, B)
 * receives cut(x_b) or E(x_b); *best_index the first maximum.  Engine
 * PQI_ENGINE_AUTO: exact for int8 and fp64 weights; for fp32 weights the
 * split-BF16 engine, whose values are approximate, within
 * 2e-5 * sum_{i<j} |w_ij| of the FP64 value (each weight keeps 2^-18 |w| after
 * the hi/lo split, the rest is FP32 accumulation); *best_index and
 * *best_value come from FP64 re-scores of the 16 best.  PQI_ENGINE_EXACT gives
 * FP64 values for every candidate. */
PQI_API int pqi_eval_batch(pqi_graph* g, const uint32_t* bits, int64_t B, int is_spin_glass,
                   double* values, int64_t* best_index, double* best_value);

/* pqi_eval_batch with an explicit engine (PQI_ENGINE_*).  With the split-BF16
 * engine, values are approximate; the 16 best by approximate value are
 * re-scored in FP64 on the device, best_index/best_value come from those FP64
 * values, and *max_rel_err (optional) receives the largest relative difference
 * between the approximate and the FP64 value among them. */
PQI_API int pqi_eval_batch_ex(pqi_graph* g, const uint32_t* bits, int64_t B, int is_spin_glass, int engine,
                              double* values, int64_t* best_index, double* best_value, double* max_rel_err);

/* pqi_eval_batch_ex with PQI_ENGINE_EXACT that allocates no device memory.
 * No counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Staging, row totals, partial sums, values and the argmax live
 * in one static device array of 32 MiB per GPU, held from the first context
 * on that GPU until the process ends.  A request whose layout needs more
 * (about B * n / 4 bytes for dense graphs, more for real weights) fails with
 * PQI_ERR_INVALID before any device work.  Results equal pqi_eval_batch_ex
 * with the exact engine.  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (other contexts, other threads) race. */
PQI_API int pqi_eval_batch_noalloc(pqi_graph* g, const uint32_t* bits, int64_t B, int is_spin_glass,
                                   double* values, int64_t* best_index, double* best_value);

/* Candidate sampler.  Replaces sample_mag + local_repulsion_choice
 * (pyqrackising/maxcut_tfim.py:36-75, 139-145; maxcut_tfim_util.py:690-724):
 * shot s = shot_offset + i draws its Hamming weight from cum_prob (ncum
 * entries, the maxcut_hamming_cdf output) and picks nodes with neighbour
 * repulsion repulsion_base^(-w).  The random stream is counter-based in
 * (seed, s, draw) and therefore reproducible; the reference's is not
 * seedable.  out: shots x n bytes, 1 = selected. */
PQI_API int pqi_sample_candidates(pqi_graph* g, const double* cum_prob, int64_t ncum, double repulsion_base,
                                  uint64_t seed, int64_t shot_offset, int64_t shots, uint8_t* out);

/* One sampling round of maxcut_tfim's accelerator path
 * (pyqrackising/maxcut_tfim.py:148-186): shots candidates sampled, scored
 * (cut, or energy in spin-glass mode) and reduced to the first best, all on
 * the device; only the best bitstring (n bytes) is copied back. */
PQI_API int pqi_sample_round(pqi_graph* g, const double* cum_prob, int64_t ncum, double repulsion_base,
                             uint64_t seed, int64_t shot_offset, int64_t shots, int is_spin_glass,
                             int64_t* best_index, double* best_value, uint8_t* best_bits);

/* Single-flip phase.  Replaces run_single_bit_flips / run_bit_flips_opencl(False)
 * (pyqrackising/spin_glass_solver.py:37-52, 156-241; kernels.cl:100-148).
 * *gain is the best cut gain, x2 in spin-glass mode; *k its first argmax. */
PQI_API int pqi_single_flip_best(pqi_graph* g, const uint8_t* theta, int is_spin_glass, int64_t* k,
                         double* gain);

/* Single-flip phase loop, resident on the device: the `while improved` loop of
 * pyqrackising/spin_glass_solver.py:604-624 restricted to its single-flip
 * branch -- take the first-argmax single flip while its gain is positive, at
 * most max_steps times.  theta (0/1 bytes) is updated in place; *accepts is
 * the number of flips applied, *gain their summed cut gain (x2 in spin-glass
 * mode, as the reference's max_energy increments). */
PQI_API int pqi_single_flip_descent(pqi_graph* g, uint8_t* theta, int is_spin_glass, int64_t max_steps,
                                    int64_t* accepts, double* gain);

/* Double-flip phase.  Replaces run_double_bit_flips / run_bit_flips_opencl(True)
 * (pyqrackising/spin_glass_solver.py:55-92, 156-241; kernels.cl:150-231). */
PQI_API int pqi_double_flip_best(pqi_graph* g, const uint8_t* theta, int is_spin_glass,
                         int64_t thread_count, int mode, int64_t* k, int64_t* l, double* gain);

/* Gray-code phase.  Replaces pick_gray_seeds + run_gray_optimization
 * (pyqrackising/spin_glass_solver.py:95-153) and run_gray_search_opencl
 * (spin_glass_solver.py:244-307; kernels.cl:877-978).
 * seed_gain: scaled cut difference of the best seed; if > 0 the best seed is
 * written to state_out and the walk is not run (spin_glass_solver.py:655-661).
 * Otherwise walk_gain is the best walker's scaled total and state_out its state. */
PQI_API int pqi_gray_search(pqi_graph* g, const uint8_t* theta, const pqi_search_params* p,
                    uint8_t* state_out, double* seed_gain, double* walk_gain);

/* The whole spin_glass_solver loop (pyqrackising/spin_glass_solver.py:603-727)
 * run on the device from theta (in: best_guess, out: final state). */
PQI_API int pqi_local_search(pqi_graph* g, uint8_t* theta, const pqi_search_params* p,
                     pqi_search_stats* stats);

/* ---- many instances at once ------------------------------------------------
 * Candidate scoring for `count` dense int8 instances of uneven sizes in one
 * persistent launch captured in a hipGraph (zero, grouped MFMA tiles, per-
 * instance first-best).  Replaces one run_cut_opencl call per instance
 * (pyqrackising/maxcut_tfim.py:212-285). */
PQI_API int pqi_batch_create(pqi_ctx* ctx, pqi_graph* const* graphs, int64_t count, int64_t cands,
                             pqi_batch** out);
PQI_API int pqi_batch_free(pqi_batch* b);
/* bits: for each instance in order, cands x ceil(n_i/32) uint32 (LSB-first). */
PQI_API int pqi_batch_set_candidates(pqi_batch* b, const uint32_t* bits);
/* best_index/best_value: one per instance; values (optional): cands per instance, in order. */
PQI_API int pqi_batch_eval(pqi_batch* b, int is_spin_glass, int64_t* best_index, double* best_value,
                           double* values);
/* Times every tile variant (contraction depth x persistent grid) `reps` times,
 * keeps the fastest and re-captures the graph.  ms_per_variant (optional). */
PQI_API int pqi_batch_autotune(pqi_batch* b, int reps, int* chosen, double* ms_per_variant, int max_variants);
PQI_API int pqi_batch_set_variant(pqi_batch* b, int variant);
PQI_API int pqi_batch_info(const pqi_batch* b, int64_t* ntiles, int64_t* algorithmic_ops, int64_t* executed_ops,
                           int64_t* weight_bytes, int* variant);

/* The spin_glass_solver loop on `count` instances at once (all of one weight
 * kind, any sizes): every phase is one launch over the instances that take
 * that branch of the reference loop.  Each instance makes exactly the moves
 * pqi_local_search makes on it alone.  thetas[i] (n_i bytes) in: start, out:
 * final state; stats: count records.  gray_iterations < 0 means n_i^2 per
 * instance.  Only the reference double-flip mode. */
PQI_API int pqi_batch_local_search(pqi_ctx* ctx, pqi_graph* const* graphs, int64_t count, uint8_t* const* thetas,
                                   const pqi_search_params* p, pqi_search_stats* stats);

/* ---- flip-proposal network --------------------------------------------------
 * No counterpart in the reference: a 3-layer message-passing net (hidden 64)
 * that scores which spins to flip; every proposal it makes is checked by the
 * exact gain kernels.  Parameters are FP32 masters; layer transforms run on
 * BF16 MFMA (precision 1, default) or FP32 (precision 0).  Features per node:
 * (gain / strength, degree / 16, spin, min(age, 64) / 64). */
PQI_API int pqi_mpnn_create(pqi_ctx* ctx, uint64_t seed, pqi_mpnn** out);
PQI_API int pqi_mpnn_free(pqi_mpnn* m);
PQI_API int pqi_mpnn_num_params(void);
PQI_API int pqi_mpnn_set_precision(pqi_mpnn* m, int precision);
PQI_API int pqi_mpnn_get_params(pqi_mpnn* m, float* out);
PQI_API int pqi_mpnn_set_params(pqi_mpnn* m, const float* in);
/* Concatenate the graphs' adjacencies into one block-diagonal batch. */
PQI_API int pqi_mpnn_bind(pqi_mpnn* m, pqi_graph* const* graphs, int64_t count, int64_t* total_nodes);
PQI_API int pqi_mpnn_forward(pqi_mpnn* m, const float* feats, float* scores);
PQI_API int pqi_mpnn_loss_grad(pqi_mpnn* m, const float* feats, const float* labels, double* loss, float* grad);
PQI_API int pqi_mpnn_train_step(pqi_mpnn* m, const float* feats, const float* labels, float lr, double* loss);

/* Batched search with learned proposals: up to `passes` rounds of network
 * scores -> score order -> one first-improvement pass with exact gains over the
 * spins whose score (a logit) exceeds min_score, then
 * the loop of pqi_batch_local_search.  The network must be bound to exactly
 * these graphs, in this order.  targets (optional, count values: energy in
 * spin-glass mode, cut otherwise): an instance stops once it reaches its own.
 * p->max_rounds < 0 stops after the learned phase (no round of the loop).
 * passes = 0 runs the plain batched loop (with targets, if given) and does
 * not use the network, which may then be bound to other graphs. */
PQI_API int pqi_learned_search(pqi_ctx* ctx, pqi_mpnn* net, pqi_graph* const* graphs, int64_t count,
                               uint8_t* const* thetas, const pqi_search_params* p, const double* targets,
                               int64_t passes, float min_score, pqi_search_stats* stats);

/* Exact re-score of one state: cut and E in FP64 (int64 for int8 weights). */
PQI_API int pqi_score(pqi_graph* g, const uint8_t* theta, double* cut, double* energy);

/* Belief-propagation marginals.  Replaces belief_propagation_marginals_dense
 * and belief_propagation_marginals (pyqrackising/spin_glass_solver.py:311-415;
 * spin_glass_solver_sparse.py:378-503): damped Jacobi sweeps of
 * m(a->b) <- d m(a->b) + (1-d) atanh(clip(tanh(w_ab) tanh(h), +-(1-1e-9))), h the sum
 * of the messages into a from its other neighbours,
 * over the arcs of the graph (dense: |w| > 1e-12; CSR: the nonzero entries; a
 * zero entry of the reference only ever carries zero messages from a zero start),
 * at most max_iterations sweeps (the reference's max(1, int(bp_scale n))),
 * stopping early at a fixed point (a sweep that changes no message, after
 * which every sweep repeats it).  marginals (n) = tanh(sum of the messages
 * into each node).  init (optional, init_count == number of arcs) sets the
 * starting messages in arc order (rows ascending, neighbours ascending; arc
 * (a, b) holds the message from b to a); NULL starts from zero as the
 * reference does, which is a fixed point with all marginals zero.
 * *iterations_run and *arc_count (both optional) report the sweeps run and
 * the number of arcs. */
PQI_API int pqi_bp_marginals(pqi_graph* g, int64_t max_iterations, double damping, const double* init,
                             int64_t init_count, double* marginals, int64_t* iterations_run, int64_t* arc_count);

typedef struct pqi_bnb_stats {
    int64_t nodes_explored; /* nodes popped */
    int64_t nodes_pruned;   /* nodes with ub <= incumbent + 1e-9 */
    int64_t batches;        /* node batches evaluated (one kernel launch each) */
    int64_t max_stack;      /* deepest stack reached, in nodes */
    int32_t certified;      /* 1: the stack emptied; 0: time_limit stopped the search */
    int32_t pad;
    double seconds;         /* wall time of the search loop */
    double kernel_seconds;  /* summed device time of the node-evaluation kernel */
} pqi_bnb_stats;

/* Batched branch-and-bound node evaluation.  Replaces _upper_bound_batch,
 * _eval_leaves_cut / _eval_leaves_energy and _influence_scores + argmax
 * (pyqrackising/solve_maxcut_bnb.py:48-121, 187-188; sparse
 * solve_maxcut_bnb_sparse.py:32-114): nodes is count x n int8 (-1 free,
 * 0/1 fixed).  ub[k]: exact cut of the fixed pairs plus max(w, 0) for every
 * pair with a free end; leaf[k]: the cut (or, in spin-glass mode,
 * sum_{i<j} (x_i == x_j ? -w : w)) when node k has no free variable, NaN
 * otherwise; branch[k]: the first free variable of largest sum |w| to the
 * other free variables, -1 for a leaf.  kernel_seconds (optional): device
 * time of the launch.  Dense graphs only, n <= 4096, count * n <= 2^30. */
PQI_API int pqi_bnb_eval(pqi_graph* g, const int8_t* nodes, int64_t count, int is_spin_glass, double* ub,
                         double* leaf, int32_t* branch, double* kernel_seconds);

/* Warm-started branch and bound.  Replaces _branch_and_bound
 * (pyqrackising/solve_maxcut_bnb.py:128-201; sparse
 * solve_maxcut_bnb_sparse.py:119-198): each batch of up to batch_cap nodes
 * popped from the top of the stack is evaluated as in pqi_bnb_eval; nodes
 * whose bound is <= incumbent + 1e-9 are pruned, the first best leaf replaces
 * the incumbent if larger, and every other node pushes its two children (the
 * incumbent's value of the branch variable first).  The exploration order is
 * the reference's, so with the reference's batch_cap (os.cpu_count()**2) the
 * counts and result are the reference's.
 * energy_bound = 0 prunes with ub in both modes, as the reference does; in
 * spin-glass mode that bound does not bound the leaf value E = 2 cut -
 * sum_{i<j} w, so the optimum can be pruned.  energy_bound = 1 prunes
 * spin-glass nodes with 2 ub - sum_{i<j} w instead, which is exact.
 * best_bits (n bytes, 0/1) and *best_value carry the incumbent in and the
 * result out.  time_limit in seconds, negative for none; when it stops the
 * search, stats->certified is 0.  Dense graphs only, n <= 4096,
 * batch_cap * n <= 2^30. */
PQI_API int pqi_bnb(pqi_graph* g, uint8_t* best_bits, double* best_value, int is_spin_glass, int energy_bound,
                    int64_t batch_cap, double time_limit, pqi_bnb_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* PQI_H */
