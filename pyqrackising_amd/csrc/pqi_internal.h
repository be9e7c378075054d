// Internal declarations shared by the libpqi translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <cstddef>

#include <string>
#include <vector>

#include "../../include/pqi.h"

namespace pqi {

// Device-side state of one problem instance.
struct Graph {
    pqi_ctx* ctx = nullptr;
    int64_t n = 0;
    int kind = PQI_WEIGHTS_I8;  // PQI_WEIGHTS_*
    // Dense weights, n_pad x n_pad row-major, zero padded (nullptr for CSR input).
    // int8_t for PQI_WEIGHTS_I8, float for F32, double for F64.
    void* W = nullptr;
    int64_t n_pad = 0;  // multiple of 128
    // Full symmetric adjacency of the nonzeros, sorted columns per row.
    int32_t* adj_ptr = nullptr;  // n+1
    int32_t* adj_idx = nullptr;  // nnz
    void* adj_w = nullptr;       // nnz, same element type as W
    int64_t nnz = 0;
    int64_t max_degree = 0;
    int64_t max_strength = -1;   // int8 weights: max_v sum_u |w_vu| (exact); -1 otherwise
    bool unit_abs = false;       // int8 weights whose nonzeros are all +-1 (strength = degree)
    int degree_parity = -1;      // 0: every degree even, 1: every degree odd, -1: mixed
    // int8 weights with adjacency lists: fixed-stride table, row v at v * stride,
    // entries (w << 24) | u, zero padded; stride = max degree rounded up so that
    // n * stride is a multiple of 4 (nullptr when not built)
    int32_t* key_adj = nullptr;
    int64_t key_adj_stride = 0;
    double w_up = 0.0;           // sum_{i<j} w_ij
    int64_t w_up_i = 0;          // same, exact, for int8 weights
    bool use_adj = true;         // flips update fields through the adjacency lists
    void* Wbf = nullptr;         // fp32 weights only: [W_hi | W_lo] bf16, n_pad x 2 n_pad
    int64_t device_bytes = 0;
};

// Scratch buffer owned by a context, grown on demand.
struct Buffer {
    void* p = nullptr;
    size_t bytes = 0;
};

// A host thread that runs one job at a time for a context (the learned
// search's fallback, beside the caller's thread); defined in pqi_bsearch.hip.
struct Worker;
void worker_stop(Worker* w);

// Pinned host memory with its device-visible address.
struct Pinned {
    void* p = nullptr;
    void* dev = nullptr;
    size_t bytes = 0;
};

}  // namespace pqi

namespace pqi {
// Captured segments of one local-search iteration, reused while every
// pointer and size they were captured with (the key) is unchanged.
struct LsGraphs {
    std::vector<int64_t> key;
    hipGraphExec_t seg[2] = {nullptr, nullptr};
    bool failed = false;  // capture is not available: launch directly
    // the iteration (without the descent when kernels are timed) and the
    // read-back of its control record into pinned host memory, as one graph
    // (nullptr: not available)
    hipGraphExec_t iter = nullptr;
    hipGraphExec_t iter2 = nullptr;  // int8 fast path: the same iteration reporting into the second slot
    void* host_ctl = nullptr;        // pinned, sizeof(LsCtl)
    // int8 fast path: iteration reports are numbered by a device counter
    // (seq_dev, incremented by every report kernel); seq_host counts the
    // iterations queued, and is re-read from the device after a call that
    // did not finish normally (seq_dirty)
    unsigned long long* seq_dev = nullptr;
    unsigned long long seq_host = 0;
    bool seq_dirty = true;
};
}  // namespace pqi

struct pqi_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cus = 0;
    std::vector<pqi::Buffer> bufs;  // indexed by slot
    // pinned host staging, grown on demand: [0] inputs and read-backs, [1] the
    // local search's iteration reports (kept apart so that a report still
    // being written by the previous call's last, idle iteration never lands on
    // the next call's input)
    // [2] the batched search's state copies, phase records and lists
    // [3] the same for a second batched search running beside the first
    //     (the learned search's fallback, on stream2)
    pqi::Pinned pin[4];
    hipStream_t stream2 = nullptr;  // created on first use, default priority
    pqi::Worker* worker = nullptr;  // the thread that drives stream2 (created on first use)
    pqi::LsGraphs ls;
};

struct pqi_graph : pqi::Graph {};
struct pqi_mpnn;

namespace pqi {

// error plumbing
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

// ---- kernels (launchers); all enqueue on ctx->stream -------------------------

// h = W s for the state x (uint8 per node).  FT = int32 for int8 weights, double otherwise.
void launch_field(const Graph& g, const uint8_t* x, void* h, hipStream_t s);
// Host rows of `width` bytes (contiguous) -> device rows at pitch dpitch,
// the bytes past width zeroed: one host copy into the context's staging
// buffer, then a restride kernel on the device (a pitched copy from pageable
// host memory would move one row per transfer).  Enqueued on the context
// stream; the staging buffer is reused by the next call.
int upload_rows(pqi_ctx* c, void* dst, size_t dpitch, const void* src, size_t width, size_t rows);
// dst row r (dwords words) = src row r (swords <= dwords words) followed by zeros.
void launch_restride_words(uint32_t* dst, int64_t dwords, const uint32_t* src, int64_t swords, int64_t rows,
                           hipStream_t s);
// Takes the per-device address of the no-allocation entry point's scratch
// array (pqi_noalloc.hip); called by pqi_create with the device current.
int noalloc_init(int device);
// Coherent pinned host staging c->pin[which] of at least `bytes`, grown on
// demand (the stream is synchronised before an old buffer is freed).
int ensure_pinned(pqi_ctx* c, int which, size_t bytes);
// The batched search's device arena, kept by the context and grown on demand.
// set 0: the context stream's arena; set 1: the second stream's
int batch_arena(pqi_ctx* c, int set, size_t bytes, void** out);
// Drift check of an incrementally maintained fp64 field: h is recomputed
// exactly from W and x and overwritten; entries that had moved by more than
// 1e-9 max(1, |h_exact|) are counted.  Skipped when gate is non-null and
// gate->again is set (a descent continuing in the same round).
struct FieldRepair {
    unsigned long long checks;    // recomputes run
    unsigned long long entries;   // entries beyond the tolerance
    unsigned long long max_bits;  // largest |h - h_exact| (bits of a non-negative double)
};
struct LsCtl;
void launch_field_repair(const Graph& g, const uint8_t* x, void* h, const LsCtl* gate, FieldRepair* fr,
                         hipStream_t s);
// Test hook: adds 1e-3 ((v mod 7) - 3) to every fp64 field entry.
void launch_perturb_field(const Graph& g, void* h, hipStream_t s);

// Exact score of the states x[b] (uint8, row stride n): out_e[b] = E(x_b) = -(1/2) s^T W s,
// out_cut[b] = cut(x_b); FP64 (int64 for int8 weights).  hbuf holds one field.
void launch_score_with(const Graph& g, const uint8_t* x, int64_t count, double* out_cut, double* out_e,
                       void* hbuf, hipStream_t s);

// Batch energies of packed candidates: q[b] = s_b^T W s_b (doubles; exact for int8).
// bits: B rows of words_per_row uint32 (words_per_row >= n_pad/32 for dense graphs).
// Dense graphs use MFMA (int8) or FP64 tiles; CSR graphs the LDS-staged CSR kernel.
void launch_batch_quadratic(const Graph& g, const uint32_t* bits, int64_t words_per_row, int64_t B,
                            double* q, void* scratch, hipStream_t s);
size_t batch_scratch_bytes(const Graph& g, int64_t B);
// int8 MFMA contraction only: qi[b] += s_b^T W s_b (int64), qi zeroed by the caller.
void launch_batch_i8_raw(const Graph& g, const uint32_t* bits, int64_t wpr, int64_t B, unsigned long long* qi,
                         hipStream_t s, const int32_t* gate = nullptr);
// Split-BF16 MFMA energies for fp32 weights (needs g.Wbf); q approximate (see pqi_eval.hip).
void launch_batch_bf16(const Graph& g, const uint32_t* bits, int64_t wpr, int64_t B, double* q, void* scratch,
                       hipStream_t s);
size_t batch_bf16_scratch_bytes(const Graph& g, int64_t B);
void launch_split_bf16(const float* W, int64_t n_pad, uint16_t* Wcat, hipStream_t s);
// out[r] = bits row sel[r], r < k.
void launch_gather_rows(const uint32_t* bits, int64_t wpr, const int64_t* sel, int64_t k, uint32_t* out,
                        hipStream_t s);

// One-work-group phase kernels over the state (x, h) of a graph.
struct PhaseOut {
    double gain;      // unscaled cut gain of the selected move
    int64_t k, l;     // selected nodes (l = -1 for a single flip)
    int64_t accepts;  // moves applied inside the kernel
    int64_t evals;    // flip evaluations performed
    double gain_sum;  // sum of applied gains
};

// Repeated single flips (first argmax, strictly positive gain) until none improves.
// With events e0/e1, they are recorded at the start and end of the kernel itself.
// With a gate, the launch does nothing while *gate is nonzero.
void launch_descent(const Graph& g, uint8_t* x, void* h, int64_t max_steps, PhaseOut* out, hipStream_t s,
                    hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr, const int32_t* gate = nullptr);
// Best single flip only (no move applied).
void launch_single_best(const Graph& g, const uint8_t* x, const void* h, PhaseOut* out, hipStream_t s);
// Best double flip: reference semantics (last pair of each of thread_count threads) or exhaustive.
void launch_double_best(const Graph& g, const uint8_t* x, const void* h, int64_t thread_count, int mode,
                        PhaseOut* out, void* scratch, hipStream_t s);
size_t double_scratch_bytes(const Graph& g);
// Apply flips of nodes k (and l if >= 0) to (x, h).
void launch_flip(const Graph& g, uint8_t* x, void* h, int64_t k, int64_t l, hipStream_t s);

// Gray seeds: packed candidate rows for seeds [0, count) of theta.
// qzero / rzero (optional): row totals and rank counts set to zero
void launch_seed_bits(const uint8_t* x, int64_t n, int64_t words_per_row, int64_t block_size,
                      int64_t count, uint32_t* bits, hipStream_t s, unsigned long long* qzero = nullptr,
                      const int32_t* gate = nullptr, int32_t* rzero = nullptr);
// int8 MFMA seeds: stable top k by row total (smallest q first = largest cut difference,
// ties by index) into sel (rank counts zeroed beforehand)
void launch_topk_q(const unsigned long long* qi, int64_t count, int64_t k, int32_t* rank, int64_t* sel,
                   hipStream_t s);
// the ranking step alone (the top k then taken by k_ls_seed_take)
void launch_rank_q(const unsigned long long* qi, int64_t count, int32_t* rank, hipStream_t s);
// Seed cut differences from int64 quadratic forms (int8 weights): q_base = s^T h of
// (x, h), delta[b] = (q_base - qi[b]) / 4; also clears rank[0..count).
void launch_seed_delta_i8(const uint8_t* x, const int32_t* h, int64_t n, const unsigned long long* qi,
                          int64_t count, double* delta, int32_t* rank, hipStream_t s);
// Stable top-k of v by rank counting with the ranks already cleared, then
// sel[r] = index of rank r and seld[r] = its value, r < k.
void launch_topk_gather(const double* v, int64_t count, int64_t k, int64_t* sel, double* seld, int32_t* rank,
                        hipStream_t s);
// Seed cut differences from the field (low-rank formula).
void launch_seed_lowrank(const Graph& g, const uint8_t* x, const void* h, int64_t block_size, int64_t count,
                         double* delta, hipStream_t s);
// Convert quadratic forms to cut differences: delta[b] = (q_base - q[b]) / 4.
void launch_q_to_delta(const double* q, const double* q_base, int64_t count, double* delta, hipStream_t s);
// Stable top-k (value descending, index ascending): sel[r] = index of rank r, for r < k.
void launch_topk(const double* v, int64_t count, int64_t k, int64_t* sel, void* scratch, hipStream_t s);
size_t topk_scratch_bytes(int64_t count);
// out[i] = v[sel[i]] for i < k.
void launch_gather(const double* v, const int64_t* sel, int64_t k, double* out, hipStream_t s);

// Gray walk: walker w starts from seed sel[w] of (x, h) and walks; writes its
// final state/field to wx/wh (row stride n), its total cut difference to we[w]
// and the number of gain evaluations it computed to wevals[w].  With a gate,
// the launch does nothing when *gate != 0 at the time it runs.
void launch_gray_walk(const Graph& g, const uint8_t* x, const void* h, const int64_t* sel,
                      const double* sel_delta, int64_t block_size, int64_t walkers,
                      int64_t thread_iterations, uint8_t* wx, void* wh, double* we, int64_t* wevals,
                      hipStream_t s, const int32_t* gate = nullptr);

// ---- device-resident control of the single-instance search loop ----------------
// One iteration of the loop (descent, double flip, Gray seeds, Gray walk,
// reheat) is enqueued as a whole; small control kernels read each phase's
// result on the device, accept or reject its move and skip the rest of the
// iteration once a move was accepted.  The host synchronises once per
// iteration.  Every accumulation happens in the order of the reference loop.
// A new round starts after this iteration (gate of the field repair, which
// the iteration graph runs at its end for the next iteration): the search
// goes on, the descent is not being continued, and the round bound allows it.
struct LsCtl;
__host__ __device__ inline bool ls_new_round(const LsCtl* c);
__host__ __device__ inline bool ls_stopped(const LsCtl* c);

struct LsCtl {
    int32_t restart;    // a move of this iteration was accepted: later phases skip
    int32_t again;      // the descent hit its step bound: the next iteration continues it
    int32_t done;       // the search ended
    int32_t skip_seed;  // the best seed is not taken
    int32_t seeds_ran, walk_ran;
    int32_t phase, pad;
    int64_t rounds;
    int64_t max_rounds;  // the caller's bound (0: none); set once
    double max_energy, scale, reheat, top;
    int64_t single_evals, single_accepts, double_evals, double_accepts, seed_evals, seed_accepts, walk_accepts,
        reheat_accepts, gray_evals, gray_evals_computed, descent_evals;
    uint64_t mark;      // device wall clock at the last phase change
    uint64_t ticks[5];  // per phase: single, double, seeds, walk, reheat
};

// The search has stopped: done, or the round bound reached outside a continued descent.
__host__ __device__ inline bool ls_stopped(const LsCtl* c) {
    return c->done || (!c->again && c->max_rounds > 0 && c->rounds >= c->max_rounds);
}

__host__ __device__ inline bool ls_new_round(const LsCtl* c) { return !c->again && !ls_stopped(c); }
void launch_ls_begin(LsCtl* c, hipStream_t s);
// What the host reads after each iteration of the int8 fast path (pinned host memory).
struct LsReport {
    LsCtl ctl;
    FieldRepair fix;
    double cut, energy;  // exact score of the state after the iteration
    unsigned long long check;  // ls_report_check of this report, its number and its state
    unsigned long long seq;  // written last: the number of the report
};
// Check word of a report: its number, its 8-byte words before `check` (control
// record, repairs, score; the structs have no padding) and the sum of its
// state's terms, each word weighted by an odd multiplier of its position, so
// that a record left by another report (another number) or read in part fails it.
__host__ __device__ inline unsigned long long ls_state_term(unsigned long long word, unsigned long long j) {
    return (word ^ j) * (0xA24BAED4963EE407ull + 2ull * j);
}
__host__ __device__ inline unsigned long long ls_report_check(const LsReport& r, unsigned long long xsum,
                                                              unsigned long long k) {
    const unsigned long long* w = (const unsigned long long*)&r;
    unsigned long long h = xsum + k * 0x9E3779B97F4A7C15ull;
    for (unsigned long long i = 0; i < offsetof(LsReport, check) / 8; ++i)
        h += (w[i] ^ i) * (0xD6E8FEB86659FD93ull + 2ull * i);
    return h;
}
// int8 weights with adjacency lists: the whole start of a local search in one launch
// (theta_dev: the caller's 0/1 bytes in pinned host memory, device-visible address)
void launch_ls_start_i8(const Graph& g, const uint8_t* theta_dev, uint8_t* x, void* h, uint8_t* rx, void* rh,
                        LsCtl* c, FieldRepair* fix, double scale, int64_t max_rounds, int is_spin_glass,
                        hipStream_t s);
// int8 weights: control record, repair counters, exact score and state into a
// pinned report numbered from *seq, then the next iteration's control reset
void launch_ls_report_i8(const Graph& g, const uint8_t* x, const void* h, LsCtl* c, const FieldRepair* fix,
                         LsReport* rep_dev, uint8_t* x_out_dev, unsigned long long* seq, hipStream_t s);
// max_energy = score[1] (E, spin glass) or score[0] (cut): the exact score of the start
void launch_ls_start_energy(LsCtl* c, const double* score, int is_spin_glass, hipStream_t s);
void launch_ls_after_descent(LsCtl* c, const PhaseOut* o, int64_t cap, hipStream_t s);
void launch_ls_after_double(const Graph& g, LsCtl* c, const PhaseOut* o, uint8_t* x, void* h, int64_t evals,
                            hipStream_t s);
// Fused control steps (reference-mode double flips): descent bookkeeping + double flip +
// its decision; Gray-seed decision + adoption of the best seed by in-place mutation;
// walk decision + post-reheat single and double flips and their decisions.
// Gray-seed rows written by the extra work groups of k_ls_descent_double
// (count = 0: none); qzero and rzero are the row totals and rank counts to clear.
struct SeedRows {
    int64_t wpr = 0, bs = 1, count = 0;
    uint32_t* bits = nullptr;
    unsigned long long* qzero = nullptr;
    int32_t* rzero = nullptr;
};
void launch_ls_descent_double(const Graph& g, LsCtl* c, const PhaseOut* dout, int64_t cap, uint8_t* x, void* h,
                              int64_t thread_count, PhaseOut* dout2, int64_t evals, hipStream_t s,
                              const SeedRows* sr = nullptr);
// With qi (int8 MFMA seeds), the selected seeds' cut differences are computed here into seld;
// with rank as well, the top G by rank are first placed into sel.
void launch_ls_seed_take(const Graph& g, LsCtl* c, double* seld, int64_t count, int64_t* sel, int64_t bs,
                         uint8_t* x, void* h, hipStream_t s, const unsigned long long* qi = nullptr, int64_t G = 0,
                         const int32_t* rank = nullptr);
// Where the int8 fast path's report goes (rep = nullptr: no report).
struct LsReportTo {
    LsReport* rep = nullptr;
    uint8_t* x_out = nullptr;
    unsigned long long* seq = nullptr;
    const FieldRepair* fix = nullptr;
    int64_t w_up_i = 0;
};
// With a report (int8 weights), the iteration's report and the next control reset follow.
void launch_ls_walk_reheat(const Graph& g, LsCtl* c, const double* we, const int64_t* wev, int64_t walkers,
                           int64_t gray_evals, const uint8_t* wx, const void* wh, uint8_t* x, void* h, uint8_t* rx,
                           void* rh, PhaseOut* dout, int64_t thread_count, PhaseOut* dout2, int64_t dbl_evals,
                           hipStream_t s, const LsReportTo* rt = nullptr);
void launch_ls_after_walk(const Graph& g, LsCtl* c, const double* we, const int64_t* wev, int64_t walkers,
                          int64_t gray_evals, const uint8_t* wx, const void* wh, uint8_t* x, void* h, uint8_t* rx,
                          void* rh, hipStream_t s);
void launch_ls_reheat_single(const Graph& g, LsCtl* c, const PhaseOut* o, uint8_t* x, void* h, uint8_t* rx,
                             void* rh, hipStream_t s);
void launch_ls_reheat_double(const Graph& g, LsCtl* c, const PhaseOut* o, uint8_t* x, void* h, uint8_t* rx,
                             void* rh, int64_t evals, hipStream_t s);

// q = s^T h of the state (x, h), FP64 (exact int64 arithmetic for int8 weights).
void launch_quad(const Graph& g, const uint8_t* x, const void* h, double* q, hipStream_t s);

size_t field_bytes(const Graph& g);  // bytes per field entry (4 or 8)

// Packed-key descent for int8 weights: key width (0 when it does not apply)
// and the dynamic LDS it uses (adjacency staged when it fits in `limit`).
int key_shift(const Graph& g);
size_t key_lds_bytes(const Graph& g, size_t limit);
// Registers per lane of the register-resident packed-key descent of g (0: it
// does not apply); *table_bytes receives the size of its neighbour table.
int reg_descent_plan(const Graph& g, size_t* table_bytes);
// words per lane (1 or 2) of the bit-plane descent for g, or 0 when it does not apply
int plane_descent_plan(const Graph& g, size_t* table_bytes);
// 1 when every degree of g has the parity of S (bit-plane values halved)
int plane_half(const Graph& g, int64_t S);

// ---- batched search: per-instance descriptors in device memory ----------------
struct BInstRaw {
    const void* W;  // dense weights (or nullptr)
    int64_t ld, n;
    const int32_t* ap;
    const int32_t* ai;
    const void* aw;
    int32_t use_adj;
    int32_t key_shift;     // packed-key descent width for int8 weights (0: not used)
    const int32_t* kadj;   // Graph::key_adj and its stride
    int64_t kstride;
    uint8_t* x;  // best state and field
    void* h;
    uint8_t* rx;  // reheat state and field
    void* rh;
    uint8_t* wx;  // walkers: G x n states and fields
    void* wh;
    double* we;
    int64_t* wev;
    double* delta;  // seed_count seed gains
    int32_t* rank;
    int64_t* sel;   // G selected seeds
    double* seld;
    PhaseOut* out;
    int64_t seed_count, bs, G, T;
    // learned proposals (nullptr when unused)
    float* feat;         // n x 4 features, a slice of the network's batch input
    const float* score;  // n scores, a slice of the network's batch output
    int32_t* perm;       // n: nodes in descending score order
    uint8_t* age;        // n: passes since each node last flipped (saturating)
};

// which = 0: the best state (x, h); 1: the reheat state (rx, rh)
void launch_b_field(int kind, const BInstRaw* d, const int32_t* list, int64_t L, int64_t max_n, int all_sparse,
                    hipStream_t s);
// all_sparse: every listed instance updates through adjacency lists (enables the
// two-level argmax variant); key_lds > 0: every instance has key_shift > 0 and
// the packed-key descent runs with that much dynamic LDS; reg_r > 0: every
// instance takes the register-resident descent with reg_r registers per lane,
// neighbour rows of at most reg_stride entries and neighbour tables of up to
// reg_lds bytes (staged in LDS where they fit); plane_s >= 0 (with reg_r > 0):
// every instance takes the bit-plane descent, plane_s = 2 S + H with S the
// largest row strength and H = 1 when every degree has the parity of S
void launch_b_descent(int kind, const BInstRaw* d, const int32_t* list, int64_t L, int64_t max_n, int64_t max_steps,
                      int which, int all_sparse, size_t key_lds, int reg_r, int64_t reg_stride, size_t reg_lds,
                      int plane_s, hipStream_t s,
                      hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// dynamic LDS limit of the batched packed-key descent
size_t b_key_lds_limit();
void launch_b_double(int kind, const BInstRaw* d, const int32_t* list, int64_t L, int which, hipStream_t s);
void launch_b_flip(int kind, const BInstRaw* d, const int32_t* list, int64_t L, const int64_t* moves, int which,
                   hipStream_t s);
void launch_b_seeds(int kind, const BInstRaw* d, const int32_t* list, int64_t L, int64_t max_seeds, int64_t max_G,
                    hipStream_t s);
// reg_r, reg_stride, reg_lds: as for launch_b_descent (register-resident walkers when reg_r > 0)
void launch_b_walk(int kind, const BInstRaw* d, const int32_t* list, int64_t L, int64_t max_n, int64_t max_G,
                   int64_t T_override, int64_t G_override, int reg_r, int64_t reg_stride, size_t reg_lds,
                   hipStream_t s);
// src[list position] = walker index, or -1 for the reheat state; dst 0 = best, 1 = reheat
void launch_b_copy(int kind, const BInstRaw* d, const int32_t* list, int64_t L, int64_t max_n, const int32_t* src,
                   int dst, hipStream_t s);
void launch_b_quad(int kind, const BInstRaw* d, const int32_t* list, int64_t L, double* q, hipStream_t s);
// learned proposals: node features from (x, h, age); score order; one first-improvement
// pass over the nodes in that order with exact gains (out->accepts, out->gain_sum)
void launch_b_features(int kind, const BInstRaw* d, const int32_t* list, int64_t L, int64_t max_n, hipStream_t s);
void launch_b_order(const BInstRaw* d, const int32_t* list, int64_t L, hipStream_t s);
// Learned passes without a synchronisation per pass: an instance's running
// value (start value plus scale times each pass's gain sum, in pass order),
// its target, and whether it still takes passes (its last pass moved and it
// is below its target); each pass's record goes to outs[pass].
struct LpGate {
    double energy, target, scale;
    int32_t active, pad;
    PhaseOut* outs;
};
// gate: nullptr, or one LpGate per instance (an inactive instance's launch
// returns at once; an active one records its pass and updates its gate)
void launch_b_learned_pass(int kind, const BInstRaw* d, const int32_t* list, int64_t L, int64_t max_n,
                           float min_score, LpGate* gate, int pass, hipStream_t s);

// candidate sampler (pqi_sampler.hip): shots rows of packed candidates from the
// Hamming-weight CDF and the neighbour-repulsion picker, shot indices shot0..
int launch_sampler(const Graph& g, const double* cum, int64_t ncum, double base, uint64_t seed, int64_t shot0,
                   int64_t shots, uint32_t* bits, int64_t wpr, hipStream_t s);

// network internals used by the learned search (pqi_mpnn.hip)
float* mpnn_features(pqi_mpnn* m);
const float* mpnn_scores(pqi_mpnn* m);
int64_t mpnn_node_offset(pqi_mpnn* m, int64_t i);
int64_t mpnn_total_nodes(pqi_mpnn* m);
void mpnn_forward_device(pqi_mpnn* m, hipStream_t s);

}  // namespace pqi
