// Belief-propagation warm start on the GPU
// (pyqrackising/spin_glass_solver.py:311-421 belief_propagation_marginals_dense /
// bp_warm_start_dense; spin_glass_solver_sparse.py:378-514
// belief_propagation_marginals / bp_warm_start).
//
// Messages live on directed arcs in adjacency order: rows ascending,
// neighbours ascending within a row, and arc (a, b) holds the message from b
// to a, so a node's incoming messages are contiguous.  One iteration is a
// Jacobi sweep, as in the reference (all new messages from the old ones):
//   H[a]          = sum of the messages into a
//   m(a -> b)_new = d m(a -> b) + (1 - d) atanh(clip(tanh(w_ab) tanh(H[a] - m(b -> a))))
// with clip to [-1 + 1e-9, 1 - 1e-9]; H[a] - m(b -> a) is the reference's sum
// over a's other neighbours.  m(a -> b) sits at arc (b, a), found through
// rev[].  The marginals are tanh(H).
//
// The sweep is a deterministic function of the message state, so once a sweep
// changes no message every later one repeats it: the loop stops there with
// the state, and so the marginals, that the full iteration count would give.
// The reference starts all messages at zero, which is such a fixed point
// (tanh(0) = 0): its marginals are all zero and its warm start is the
// all-zeros partition, reached here after one sweep.
//
// Dense graphs take the arcs with |w| > 1e-12 (the dense reference's filter);
// graphs built from CSR take the nonzero entries (the sparse reference takes
// every stored entry; from its zero start a zero-weight entry only ever carries
// zero messages, which add nothing).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "pqi_internal.h"

namespace pqi {

namespace {

constexpr double kBpThreshold = 1e-12;
constexpr int kBpWaves = 4;

__device__ __forceinline__ double weight_of(const void* p, int kind, int64_t e) {
    switch (kind) {
        case PQI_WEIGHTS_I8: return (double)((const int8_t*)p)[e];
        case PQI_WEIGHTS_F32: return (double)((const float*)p)[e];
        default: return ((const double*)p)[e];
    }
}

__device__ __forceinline__ int row_of_wave() { return blockIdx.x * kBpWaves + (threadIdx.x >> 6); }

// dense weights: number of arcs of each row
__global__ __launch_bounds__(64 * kBpWaves) void k_bp_count(const void* W, int kind, int64_t ld, int n,
                                                            int32_t* cnt) {
    const int lane = threadIdx.x & 63, a = row_of_wave();
    if (a >= n) return;
    int c = 0;
    for (int b0 = 0; b0 < n; b0 += 64) {
        const int b = b0 + lane;
        const bool live = b < n && b != a && fabs(weight_of(W, kind, (int64_t)a * ld + b)) > kBpThreshold;
        c += __popcll(__ballot(live));
    }
    if (lane == 0) cnt[a] = c;
}

// ptr[0] = 0, ptr[i + 1] = cnt[0] + ... + cnt[i]; one work group of 1024
__global__ __launch_bounds__(1024) void k_bp_scan(const int32_t* cnt, int n, int64_t* ptr) {
    __shared__ int64_t wsum[16];
    __shared__ int64_t carry;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (threadIdx.x == 0) {
        carry = 0;
        ptr[0] = 0;
    }
    __syncthreads();
    for (int base = 0; base < n; base += 1024) {
        const int i = base + threadIdx.x;
        int64_t v = i < n ? cnt[i] : 0;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int64_t o = __shfl_up(v, d, 64);
            if (lane >= d) v += o;
        }
        if (lane == 63) wsum[wid] = v;
        __syncthreads();
        int64_t off = carry;
        for (int w = 0; w < wid; ++w) off += wsum[w];
        if (i < n) ptr[i + 1] = off + v;
        __syncthreads();
        if (threadIdx.x == 1023) carry = off + v;
        __syncthreads();
    }
}

// dense weights: the arcs of each row in column order, and tanh(w)
__global__ __launch_bounds__(64 * kBpWaves) void k_bp_fill_dense(const void* W, int kind, int64_t ld, int n,
                                                                 const int64_t* ptr, int32_t* idx, double* tw) {
    const int lane = threadIdx.x & 63, a = row_of_wave();
    if (a >= n) return;
    int64_t p = ptr[a];
    for (int b0 = 0; b0 < n; b0 += 64) {
        const int b = b0 + lane;
        const double w = b < n ? weight_of(W, kind, (int64_t)a * ld + b) : 0.0;
        const bool live = b < n && b != a && fabs(w) > kBpThreshold;
        const uint64_t m = __ballot(live);
        if (live) {
            const int64_t q = p + __popcll(m & ((1ull << lane) - 1ull));
            idx[q] = b;
            tw[q] = tanh(w);
        }
        p += __popcll(m);
    }
}

// adjacency lists: int64 row pointers and tanh(w)
__global__ void k_bp_from_adj(const int32_t* aptr, const void* aw, int kind, int n, int64_t nnz, int64_t* ptr,
                              double* tw) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t <= n) ptr[t] = aptr[t];
    if (t < nnz) tw[t] = tanh(weight_of(aw, kind, t));
}

// rev[e] for arc e = (a, b): the position of arc (b, a) (binary search in row b)
__global__ __launch_bounds__(64 * kBpWaves) void k_bp_rev(const int64_t* ptr, const int32_t* idx, int n,
                                                          int32_t* rev) {
    const int lane = threadIdx.x & 63, a = row_of_wave();
    if (a >= n) return;
    for (int64_t e = ptr[a] + lane; e < ptr[a + 1]; e += 64) {
        const int b = idx[e];
        int64_t lo = ptr[b], hi = ptr[b + 1];
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (idx[mid] < a) lo = mid + 1;
            else hi = mid;
        }
        rev[e] = (int32_t)lo;  // the lists are symmetric: idx[lo] == a
    }
}

// H[a] = sum of the messages into a; tanh(H) when marginal is set
__global__ __launch_bounds__(64 * kBpWaves) void k_bp_field(const int64_t* ptr, const double* in, int n, double* H,
                                                            int marginal) {
    const int lane = threadIdx.x & 63, a = row_of_wave();
    if (a >= n) return;
    double s = 0.0;
    for (int64_t e = ptr[a] + lane; e < ptr[a + 1]; e += 64) s += in[e];
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) s += __shfl_xor(s, m, 64);
    if (lane == 0) H[a] = marginal ? tanh(s) : s;
}

__global__ __launch_bounds__(64 * kBpWaves) void k_bp_sweep(const int64_t* ptr, const int32_t* rev,
                                                            const double* tw, const double* in_old,
                                                            double* in_new, const double* H, int n, double damping,
                                                            int* changed) {
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63, a = row_of_wave();
    if (a >= n) return;
    const double lo = -1.0 + 1e-9, hi = 1.0 - 1e-9;
    bool any = false;
    for (int64_t e = ptr[a] + lane; e < ptr[a + 1]; e += 64) {
        double p = tw[e] * tanh(H[a] - in_old[e]);
        p = p < lo ? lo : (p > hi ? hi : p);  // NaN passes, as np.clip
        const int32_t r = rev[e];
        const double old = in_old[r];
        const double m = damping * old + (1.0 - damping) * atanh(p);
        in_new[r] = m;
        any |= !(m == old);
    }
    if (__ballot(any) && lane == 0) *changed = 1;
}

__global__ void k_bp_zero_flag(int* f) { *f = 0; }

}  // namespace

}  // namespace pqi

using namespace pqi;

extern "C" {

int pqi_bp_marginals(pqi_graph* g, int64_t max_iterations, double damping, const double* init, int64_t init_count,
                     double* marginals, int64_t* iterations_run, int64_t* arc_count) {
    if (!g || !marginals) return fail(PQI_ERR_INVALID, "null argument");
    if (max_iterations < 1) return fail(PQI_ERR_INVALID, "max_iterations must be >= 1");
    const int64_t n = g->n;
    if (n < 1 || n > INT32_MAX / 2) return fail(PQI_ERR_UNSUPPORTED, "graph size out of range");
    pqi_ctx* c = g->ctx;
    if (hipSetDevice(c->device) != hipSuccess) return fail(PQI_ERR_HIP, "hipSetDevice failed");
    hipStream_t s = c->stream;
    const unsigned rows = (unsigned)((n + kBpWaves - 1) / kBpWaves);
    const dim3 wb(64 * kBpWaves);

    std::vector<void*> owned;
    auto release = [&]() {
        for (void* p : owned) (void)hipFree(p);
    };
    auto dalloc = [&](void** p, size_t bytes) {
        if (hipMalloc(p, std::max<size_t>(bytes, 8)) != hipSuccess) return false;
        owned.push_back(*p);
        return true;
    };
    int64_t* ptr = nullptr;
    int32_t* idx = nullptr;
    double* tw = nullptr;
    if (!dalloc((void**)&ptr, sizeof(int64_t) * (n + 1))) return fail(PQI_ERR_OOM, "device allocation failed");
    int64_t arcs = 0;
    if (g->W) {
        int32_t* cnt = nullptr;
        if (!dalloc((void**)&cnt, sizeof(int32_t) * n)) {
            release();
            return fail(PQI_ERR_OOM, "device allocation failed");
        }
        hipLaunchKernelGGL(k_bp_count, dim3(rows), wb, 0, s, g->W, g->kind, g->n_pad, (int)n, cnt);
        hipLaunchKernelGGL(k_bp_scan, dim3(1), dim3(1024), 0, s, cnt, (int)n, ptr);
        if (hipMemcpyAsync(&arcs, ptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            release();
            return fail(PQI_ERR_HIP, "counting the arcs failed");
        }
        if (arcs > INT32_MAX) {
            release();
            return fail(PQI_ERR_UNSUPPORTED, "too many arcs for belief propagation");
        }
        if (!dalloc((void**)&idx, sizeof(int32_t) * arcs) || !dalloc((void**)&tw, sizeof(double) * arcs)) {
            release();
            return fail(PQI_ERR_OOM, "device allocation failed");
        }
        hipLaunchKernelGGL(k_bp_fill_dense, dim3(rows), wb, 0, s, g->W, g->kind, g->n_pad, (int)n, ptr, idx, tw);
    } else {
        arcs = g->nnz;
        idx = g->adj_idx;
        if (!dalloc((void**)&tw, sizeof(double) * arcs)) {
            release();
            return fail(PQI_ERR_OOM, "device allocation failed");
        }
        const int64_t t = std::max<int64_t>(arcs, n + 1);
        hipLaunchKernelGGL(k_bp_from_adj, dim3((unsigned)((t + 255) / 256)), dim3(256), 0, s, g->adj_ptr, g->adj_w,
                           g->kind, (int)n, arcs, ptr, tw);
    }
    if (arc_count) *arc_count = arcs;
    if (init && init_count != arcs) {
        release();
        return fail(PQI_ERR_INVALID, "init must hold one message per arc");
    }

    int32_t* rev = nullptr;
    double *m0 = nullptr, *m1 = nullptr, *H = nullptr;
    int* flag = nullptr;
    if (!dalloc((void**)&rev, sizeof(int32_t) * arcs) || !dalloc((void**)&m0, sizeof(double) * arcs) ||
        !dalloc((void**)&m1, sizeof(double) * arcs) || !dalloc((void**)&H, sizeof(double) * n) ||
        !dalloc((void**)&flag, sizeof(int))) {
        release();
        return fail(PQI_ERR_OOM, "device allocation failed");
    }
    hipLaunchKernelGGL(k_bp_rev, dim3(rows), wb, 0, s, ptr, idx, (int)n, rev);
    hipError_t he = init ? hipMemcpyAsync(m0, init, sizeof(double) * arcs, hipMemcpyHostToDevice, s)
                         : hipMemsetAsync(m0, 0, sizeof(double) * std::max<int64_t>(arcs, 1), s);
    if (he != hipSuccess) {
        release();
        return fail(PQI_ERR_HIP, "initialising the messages failed");
    }

    int64_t it = 0, next_check = 1;
    while (it < max_iterations) {
        hipLaunchKernelGGL(k_bp_field, dim3(rows), wb, 0, s, ptr, m0, (int)n, H, 0);
        const bool check = it + 1 == next_check || it + 1 == max_iterations;
        if (check) hipLaunchKernelGGL(k_bp_zero_flag, dim3(1), dim3(1), 0, s, flag);
        hipLaunchKernelGGL(k_bp_sweep, dim3(rows), wb, 0, s, ptr, rev, tw, m0, m1, H, (int)n, damping, flag);
        std::swap(m0, m1);
        ++it;
        if (check) {
            int hf = 1;
            if (hipMemcpyAsync(&hf, flag, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess) {
                release();
                return fail(PQI_ERR_HIP, "belief-propagation sweep failed");
            }
            if (!hf) break;  // fixed point: every later sweep repeats this state
            next_check = it + std::min<int64_t>(it, 64);
        }
    }
    hipLaunchKernelGGL(k_bp_field, dim3(rows), wb, 0, s, ptr, m0, (int)n, H, 1);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(marginals, H, sizeof(double) * n, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        release();
        return fail(PQI_ERR_HIP, "belief propagation failed");
    }
    release();
    if (iterations_run) *iterations_run = it;
    return PQI_OK;
}

}  // extern "C"
