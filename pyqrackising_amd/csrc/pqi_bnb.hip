// Branch and bound with the node evaluation on the GPU
// (pyqrackising/solve_maxcut_bnb.py: _upper_bound_batch :48-75,
// _eval_leaves_cut :78-89, _eval_leaves_energy :92-103, _influence_scores
// :106-121, the loop _branch_and_bound :128-201; the sparse file
// solve_maxcut_bnb_sparse.py:32-198 computes the same quantities from CSR).
//
// A node is an int8 row over the variables: -1 free, 0/1 fixed.  For every
// node of a batch one wavefront computes, over the symmetric dense weights,
//   ub     = sum_{i<j} (i, j fixed ? w [x_i != x_j] : max(w, 0))
//   leaf   = the cut sum_{i<j} w [x_i != x_j], or the energy
//            sum_{i<j} (x_i == x_j ? -w : w) in spin-glass mode
//   branch = the free variable with the largest sum of |w| to the other free
//            variables, the first one on ties (numpy argmax)
//   nfree  = the number of free variables.
// Lanes own columns (coalesced row reads); the influence score of column c
// accumulates down the rows in row order, the reference's own order.
//
// The host loop keeps the reference's stack discipline: pop up to batch_cap
// nodes from the top (the k-th pop is slot k), prune at ub <= best + 1e-9,
// adopt the first best leaf if it beats the incumbent, then push each
// surviving interior node's two children, the incumbent's value of the
// branch variable first.  The explored and pruned counts and the result are
// therefore the reference's for the same batch_cap.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "pqi_internal.h"

namespace pqi {

struct BnbOut {
    double ub, leaf;
    int32_t branch, nfree;
};

constexpr int kBnbWaves = 4;
constexpr int64_t kBnbMaxN = 4096;
constexpr int64_t kBnbMaxBytes = (int64_t)1 << 30;  // batch_cap * n

// nodes: bs rows of n int8.  rev = 1: stack order (the top of the stack
// last), slot s is row bs-1-s; rev = 0: slot s is row s.
template <typename WT>
__global__ __launch_bounds__(64 * kBnbWaves) void k_bnb_eval(const WT* __restrict__ W, int64_t ld, int n,
                                                             const int8_t* __restrict__ nodes, int bs, int rev,
                                                             int sg, BnbOut* __restrict__ out) {
    extern __shared__ int8_t fl[];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int n4 = (n + 3) & ~3;
    const int slot0 = blockIdx.x * kBnbWaves;
    for (int t = threadIdx.x; t < kBnbWaves * n; t += blockDim.x) {
        const int q = t / n, i = t - q * n, s = slot0 + q;
        fl[q * n4 + i] = s < bs ? nodes[(int64_t)(rev ? bs - 1 - s : s) * n + i] : (int8_t)0;
    }
    __syncthreads();
    const int slot = slot0 + wid;
    if (slot >= bs) return;  // no barrier below
    const int8_t* f = fl + wid * n4;

    double ub = 0.0, leaf = 0.0, best = -1.0;
    int nfree = 0, bk = INT_MAX;
    for (int c0 = 0; c0 < n; c0 += 64) {
        const int c = c0 + lane;
        const bool cv = c < n;
        const int fc = cv ? f[c] : 0;
        const WT* col = W + (cv ? c : 0);
        double score = 0.0;
        for (int r = 0; r < n; ++r) {
            const int fr = f[r];
            const double w = cv ? (double)col[(int64_t)r * ld] : 0.0;
            if (c > r) {
                const bool both = (fr >= 0) & (fc >= 0);
                ub += both ? (fr != fc ? w : 0.0) : (w > 0.0 ? w : 0.0);
                if (sg) leaf += fr == fc ? -w : w;
                else leaf += fr != fc ? w : 0.0;
            }
            score += (fr < 0 && r != c) ? fabs(w) : 0.0;
        }
        if (cv && fc < 0) {
            ++nfree;
            if (score > best) {  // columns ascend per lane: first maximum
                best = score;
                bk = c;
            }
        }
    }
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) {
        ub += __shfl_xor(ub, m, 64);
        leaf += __shfl_xor(leaf, m, 64);
        nfree += __shfl_xor(nfree, m, 64);
        const double ob = __shfl_xor(best, m, 64);
        const int ok = __shfl_xor(bk, m, 64);
        if (ob > best || (ob == best && ok < bk)) {
            best = ob;
            bk = ok;
        }
    }
    if (lane == 0) {
        BnbOut o;
        o.ub = ub;
        o.leaf = nfree > 0 ? __builtin_nan("") : leaf;
        o.branch = nfree > 0 ? bk : -1;
        o.nfree = nfree;
        out[slot] = o;
    }
}

static void launch_bnb_eval(const Graph& g, const int8_t* nodes, int bs, int rev, int sg, BnbOut* out,
                            hipEvent_t e0, hipEvent_t e1, hipStream_t s) {
    const int n = (int)g.n;
    const dim3 grid((unsigned)((bs + kBnbWaves - 1) / kBnbWaves)), block(64 * kBnbWaves);
    const size_t lds = (size_t)kBnbWaves * ((n + 3) & ~3);
    switch (g.kind) {
        case PQI_WEIGHTS_I8:
            hipExtLaunchKernelGGL(k_bnb_eval<int8_t>, grid, block, lds, s, e0, e1, 0, (const int8_t*)g.W, g.n_pad, n,
                                  nodes, bs, rev, sg, out);
            break;
        case PQI_WEIGHTS_F32:
            hipExtLaunchKernelGGL(k_bnb_eval<float>, grid, block, lds, s, e0, e1, 0, (const float*)g.W, g.n_pad, n,
                                  nodes, bs, rev, sg, out);
            break;
        default:
            hipExtLaunchKernelGGL(k_bnb_eval<double>, grid, block, lds, s, e0, e1, 0, (const double*)g.W, g.n_pad, n,
                                  nodes, bs, rev, sg, out);
            break;
    }
}

}  // namespace pqi

using namespace pqi;

extern "C" {

int pqi_bnb_eval(pqi_graph* g, const int8_t* nodes, int64_t count, int is_spin_glass, double* ub, double* leaf,
                 int32_t* branch, double* kernel_seconds) {
    if (!g || (count > 0 && (!nodes || !ub || !leaf || !branch))) return fail(PQI_ERR_INVALID, "null argument");
    if (!g->W) return fail(PQI_ERR_UNSUPPORTED, "branch and bound needs a dense graph");
    const int64_t n = g->n;
    if (n < 1 || n > kBnbMaxN) return fail(PQI_ERR_UNSUPPORTED, "branch and bound supports 1 <= n <= 4096");
    if (count < 0 || count > kBnbMaxBytes / n) return fail(PQI_ERR_INVALID, "count out of range");
    for (int64_t i = 0; i < count * n; ++i)
        if (nodes[i] < -1 || nodes[i] > 1) return fail(PQI_ERR_INVALID, "node entries must be -1, 0 or 1");
    if (kernel_seconds) *kernel_seconds = 0.0;
    if (count == 0) return PQI_OK;
    pqi_ctx* c = g->ctx;
    if (hipSetDevice(c->device) != hipSuccess) return fail(PQI_ERR_HIP, "hipSetDevice failed");
    hipStream_t s = c->stream;
    int8_t* d_nodes = nullptr;
    BnbOut* d_out = nullptr;
    if (hipMalloc(&d_nodes, (size_t)count * n) != hipSuccess ||
        hipMalloc(&d_out, sizeof(BnbOut) * (size_t)count) != hipSuccess) {
        if (d_nodes) (void)hipFree(d_nodes);
        return fail(PQI_ERR_OOM, "device allocation for node evaluation failed");
    }
    std::vector<BnbOut> out((size_t)count);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const bool timed = kernel_seconds && hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
    int rc = PQI_OK;
    if (hipMemcpyAsync(d_nodes, nodes, (size_t)count * n, hipMemcpyHostToDevice, s) != hipSuccess) {
        rc = fail(PQI_ERR_HIP, "uploading the nodes failed");
    } else {
        launch_bnb_eval(*g, d_nodes, (int)count, 0, is_spin_glass ? 1 : 0, d_out, timed ? e0 : nullptr,
                        timed ? e1 : nullptr, s);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(out.data(), d_out, sizeof(BnbOut) * count, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            rc = fail(PQI_ERR_HIP, "node evaluation failed");
    }
    if (rc == PQI_OK && timed) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, e0, e1) == hipSuccess) *kernel_seconds = ms * 1e-3;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(d_nodes);
    (void)hipFree(d_out);
    if (rc != PQI_OK) return rc;
    for (int64_t k = 0; k < count; ++k) {
        ub[k] = out[k].ub;
        leaf[k] = out[k].leaf;
        branch[k] = out[k].branch;
    }
    return PQI_OK;
}

int pqi_bnb(pqi_graph* g, uint8_t* best_bits, double* best_value, int is_spin_glass, int energy_bound,
            int64_t batch_cap, double time_limit, pqi_bnb_stats* stats) {
    if (!g || !best_bits || !best_value || !stats) return fail(PQI_ERR_INVALID, "null argument");
    if (!g->W) return fail(PQI_ERR_UNSUPPORTED, "branch and bound needs a dense graph");
    const int64_t n = g->n;
    if (n < 1 || n > kBnbMaxN) return fail(PQI_ERR_UNSUPPORTED, "branch and bound supports 1 <= n <= 4096");
    if (batch_cap < 1 || batch_cap > kBnbMaxBytes / n) return fail(PQI_ERR_INVALID, "batch_cap out of range");
    for (int64_t i = 0; i < n; ++i)
        if (best_bits[i] > 1) return fail(PQI_ERR_INVALID, "best_bits entries must be 0 or 1");
    pqi_ctx* c = g->ctx;
    if (hipSetDevice(c->device) != hipSuccess) return fail(PQI_ERR_HIP, "hipSetDevice failed");
    hipStream_t s = c->stream;

    int8_t* d_nodes = nullptr;
    BnbOut* d_out = nullptr;
    int8_t* h_nodes = nullptr;
    BnbOut* h_out = nullptr;
    auto release = [&]() {
        if (d_nodes) (void)hipFree(d_nodes);
        if (d_out) (void)hipFree(d_out);
        if (h_nodes) (void)hipHostFree(h_nodes);
        if (h_out) (void)hipHostFree(h_out);
    };
    if (hipMalloc(&d_nodes, (size_t)batch_cap * n) != hipSuccess ||
        hipMalloc(&d_out, sizeof(BnbOut) * (size_t)batch_cap) != hipSuccess ||
        hipHostMalloc(&h_nodes, (size_t)batch_cap * n, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&h_out, sizeof(BnbOut) * (size_t)batch_cap, hipHostMallocDefault) != hipSuccess) {
        release();
        return fail(PQI_ERR_OOM, "allocation for branch and bound failed");
    }

    std::vector<uint8_t> bits(best_bits, best_bits + n);
    double best = *best_value;
    // spin-glass mode with energy_bound: E = 2 cut - sum_{i<j} w <= 2 ub - sum_{i<j} w
    const bool ebound = is_spin_glass && energy_bound;
    const double total = g->kind == PQI_WEIGHTS_I8 ? (double)g->w_up_i : g->w_up;
    std::vector<int8_t> stack((size_t)n, (int8_t)-1);  // the root: all free
    int64_t depth = 1, explored = 0, pruned = 0, batches = 0, max_depth = 1;
    bool certified = true;
    double kernel_ms = 0.0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const bool timed = hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
    const auto t0 = std::chrono::steady_clock::now();
    int rc = PQI_OK;
    std::vector<int64_t> interior;
    while (depth > 0) {
        if (time_limit >= 0.0 &&
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > time_limit) {
            certified = false;
            break;
        }
        const int64_t bs = std::min<int64_t>(batch_cap, depth);
        std::memcpy(h_nodes, stack.data() + (size_t)(depth - bs) * n, (size_t)bs * n);
        depth -= bs;
        stack.resize((size_t)depth * n);
        explored += bs;
        ++batches;
        if (hipMemcpyAsync(d_nodes, h_nodes, (size_t)bs * n, hipMemcpyHostToDevice, s) != hipSuccess) {
            rc = fail(PQI_ERR_HIP, "uploading a batch of nodes failed");
            break;
        }
        launch_bnb_eval(*g, d_nodes, (int)bs, 1, is_spin_glass ? 1 : 0, d_out, timed ? e0 : nullptr,
                        timed ? e1 : nullptr, s);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(h_out, d_out, sizeof(BnbOut) * bs, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            rc = fail(PQI_ERR_HIP, "branch-and-bound node evaluation failed");
            break;
        }
        if (timed) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, e0, e1) == hipSuccess) kernel_ms += ms;
        }
        auto row = [&](int64_t k) { return h_nodes + (size_t)(bs - 1 - k) * n; };
        interior.clear();
        int64_t bl = -1;
        for (int64_t k = 0; k < bs; ++k) {
            const double ub = ebound ? 2.0 * h_out[k].ub - total : h_out[k].ub;
            if (ub <= best + 1e-9) {
                ++pruned;
                continue;
            }
            if (h_out[k].nfree == 0) {
                if (bl < 0 || h_out[k].leaf > h_out[bl].leaf) bl = k;
            } else {
                interior.push_back(k);
            }
        }
        if (bl >= 0 && h_out[bl].leaf > best) {
            best = h_out[bl].leaf;
            const int8_t* x = row(bl);
            for (int64_t i = 0; i < n; ++i) bits[i] = x[i] >= 1 ? 1 : 0;
        }
        for (int64_t k : interior) {
            const int32_t v = h_out[k].branch;
            const int8_t warm = (int8_t)bits[v];
            for (int8_t val : {warm, (int8_t)(1 - warm)}) {
                stack.insert(stack.end(), row(k), row(k) + n);
                stack[(size_t)depth * n + v] = val;
                ++depth;
            }
        }
        max_depth = std::max(max_depth, depth);
    }
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    release();
    if (rc != PQI_OK) return rc;
    std::memcpy(best_bits, bits.data(), (size_t)n);
    *best_value = best;
    stats->nodes_explored = explored;
    stats->nodes_pruned = pruned;
    stats->batches = batches;
    stats->max_stack = max_depth;
    stats->certified = certified ? 1 : 0;
    stats->pad = 0;
    stats->seconds = wall;
    stats->kernel_seconds = kernel_ms * 1e-3;
    return PQI_OK;
}

}  // extern "C"
