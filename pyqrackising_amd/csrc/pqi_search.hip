// Local-search kernels for gfx950: local fields, single-flip descent with
// work-group argmax, double flips, Gray seeds, top-k and Gray-code walkers.
//
// State of a search = (x, h): x[v] in {0,1} (uint8) and the local field
// h[v] = sum_u W[v][u] s_u with s = 1 - 2x.  The cut gain of flipping k is
// Delta_k = s_k h_k (pyqrackising/maxcut_tfim_util.py:334-336 computes the same
// value in O(n)); a pair gains Delta_k + Delta_l - 2 w_kl s_k s_l
// (maxcut_tfim_util.py:339-346).  Flipping k changes h_v by -2 w_vk s_k(old).
// Fields are int32 for int8 weights (exact) and FP64 otherwise.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <climits>
#include <type_traits>
#include <cstdlib>
#include <cmath>

#include "pqi_internal.h"

namespace pqi {

template <typename WT> struct FieldOf;
template <> struct FieldOf<int8_t> { using T = int32_t; };
template <> struct FieldOf<float> { using T = double; };
template <> struct FieldOf<double> { using T = double; };

template <typename T> __device__ __forceinline__ T lowest();
template <> __device__ __forceinline__ int32_t lowest<int32_t>() { return INT_MIN; }
template <> __device__ __forceinline__ int64_t lowest<int64_t>() { return LLONG_MIN; }
template <> __device__ __forceinline__ double lowest<double>() { return -INFINITY; }

// Plain-data view of a Graph for kernels.
template <typename WT>
struct GView {
    const WT* W;
    int64_t ld;
    int64_t n;
    const int32_t* ap;
    const int32_t* ai;
    const WT* aw;
    int use_adj;
    const int32_t* kadj;  // int8 weights: fixed-stride packed adjacency (or nullptr)
    int64_t kstride;
};

template <typename WT>
static GView<WT> view(const Graph& g) {
    GView<WT> v;
    v.W = (const WT*)g.W;
    v.ld = g.n_pad;
    v.n = g.n;
    v.ap = g.adj_ptr;
    v.ai = g.adj_idx;
    v.aw = (const WT*)g.adj_w;
    v.use_adj = (g.use_adj || g.W == nullptr) ? 1 : 0;
    v.kadj = g.key_adj;
    v.kstride = g.key_adj ? g.key_adj_stride : 0;
    return v;
}

// w_kl: dense lookup or binary search in the sorted adjacency row k.
template <typename WT>
__device__ __forceinline__ WT weight_of(const GView<WT>& g, int64_t k, int64_t l) {
    if (g.W) return g.W[k * g.ld + l];
    int32_t lo = g.ap[k], hi = g.ap[k + 1] - 1;
    while (lo <= hi) {
        int32_t mid = (lo + hi) >> 1;
        int32_t c = g.ai[mid];
        if (c == l) return g.aw[mid];
        if (c < l) lo = mid + 1;
        else hi = mid - 1;
    }
    return WT(0);
}

template <typename FT>
__device__ __forceinline__ FT gain_of(const uint8_t* x, const FT* h, int64_t k) {
    return x[k] ? -h[k] : h[k];
}

// (value, index) order: larger value first, then smaller index.
template <typename FT>
__device__ __forceinline__ bool better(FT av, int64_t ak, FT bv, int64_t bk) {
    return av > bv || (av == bv && ak < bk);
}

template <typename T>
__device__ __forceinline__ T shfl_xor_t(T v, int m) {
    return __shfl_xor(v, m, 64);
}
template <>
__device__ __forceinline__ int64_t shfl_xor_t<int64_t>(int64_t v, int m) {
    int lo = __shfl_xor((int)(v & 0xffffffff), m, 64);
    int hi = __shfl_xor((int)(v >> 32), m, 64);
    return ((int64_t)hi << 32) | (uint32_t)lo;
}

template <typename FT>
__device__ __forceinline__ void wave_argmax(FT& v, int64_t& k) {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) {
        FT ov = shfl_xor_t(v, m);
        int64_t ok = shfl_xor_t(k, m);
        if (better(ov, ok, v, k)) { v = ov; k = ok; }
    }
}

// Work-group argmax; every thread returns the winner.  Uses 2*nwaves LDS slots.
template <typename FT>
__device__ void block_argmax(FT& v, int64_t& k, FT* sv, int64_t* sk) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    wave_argmax(v, k);
    __syncthreads();
    if (lane == 0) { sv[wid] = v; sk[wid] = k; }
    __syncthreads();
    v = sv[0];
    k = sk[0];
    for (int i = 1; i < nw; ++i)
        if (better(sv[i], sk[i], v, k)) { v = sv[i]; k = sk[i]; }
    __syncthreads();
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) v += shfl_xor_t(v, m);
    return v;
}

// Wavefront reductions on DPP row operations (a few cycles per step instead of
// an LDS-path shuffle): quad swaps, row shifts, then row broadcasts; the
// result is complete in lane 63 and read back to every lane.
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ int dpp_mov(int v) {
    return __builtin_amdgcn_update_dpp(v, v, CTRL, ROWS, 0xF, false);
}

template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ double dpp_mov(double v) {
    const int lo = dpp_mov<CTRL, ROWS>(__double2loint(v));
    const int hi = dpp_mov<CTRL, ROWS>(__double2hiint(v));
    return __hiloint2double(hi, lo);
}

template <typename T, typename Op>
__device__ __forceinline__ T wave_reduce_dpp(T v, Op op) {
    v = op(v, dpp_mov<0xB1>(v));        // quad_perm [1,0,3,2]
    v = op(v, dpp_mov<0x4E>(v));        // quad_perm [2,3,0,1]
    v = op(v, dpp_mov<0x114>(v));       // row_shr:4
    v = op(v, dpp_mov<0x118>(v));       // row_shr:8
    v = op(v, dpp_mov<0x142, 0xA>(v));  // row_bcast:15
    v = op(v, dpp_mov<0x143, 0xC>(v));  // row_bcast:31
    if constexpr (sizeof(T) == 8) {
        const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
        const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
        return __hiloint2double(hi, lo);
    } else {
        return __builtin_amdgcn_readlane(v, 63);
    }
}

__device__ __forceinline__ int32_t max3i(int32_t a, int32_t b, int32_t c) { return max(max(a, b), c); }

// Four one-lane masks at once: m_j = (lane == l_j) ? b_j : 0 for wave-uniform
// lanes l_j.  The four compares write four scalar pairs and go out before the
// four selects, so the selects do not wait on each other (compares through
// VCC would serialise them).
__device__ __forceinline__ void lane_masks4(int lane, int l0, int l1, int l2, int l3, uint32_t b0, uint32_t b1,
                                            uint32_t b2, uint32_t b3, uint32_t& m0, uint32_t& m1, uint32_t& m2,
                                            uint32_t& m3) {
    uint64_t k0, k1, k2, k3;
    __asm__(
        "v_cmp_eq_u32_e64 %4, %8, %12\n\t"
        "v_cmp_eq_u32_e64 %5, %9, %12\n\t"
        "v_cmp_eq_u32_e64 %6, %10, %12\n\t"
        "v_cmp_eq_u32_e64 %7, %11, %12\n\t"
        "v_cndmask_b32_e64 %0, 0, %13, %4\n\t"
        "v_cndmask_b32_e64 %1, 0, %14, %5\n\t"
        "v_cndmask_b32_e64 %2, 0, %15, %6\n\t"
        "v_cndmask_b32_e64 %3, 0, %16, %7"
        : "=&v"(m0), "=&v"(m1), "=&v"(m2), "=v"(m3), "=&s"(k0), "=&s"(k1), "=&s"(k2), "=&s"(k3)
        : "s"(l0), "s"(l1), "s"(l2), "s"(l3), "v"(lane), "v"(b0), "v"(b1), "v"(b2), "v"(b3));
}

// Integer wave maximum; the identity (INT_MIN) as the DPP fill value lets
// each step fold into one v_max_i32 with a DPP source.
__device__ __forceinline__ int32_t wave_max_i32(int32_t v) {
    v = max(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
    v = max(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
    v = max(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x114, 0xF, 0xF, false));  // row_shr:4
    v = max(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x118, 0xF, 0xF, false));  // row_shr:8
    v = max(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x142, 0xA, 0xF, false));  // row_bcast:15
    v = max(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x143, 0xC, 0xF, false));  // row_bcast:31
    return __builtin_amdgcn_readlane(v, 63);
}

// First maximum over the wave: the largest value, and among the lanes holding
// it the smallest index.  Every lane returns the winner.
template <typename FT>
__device__ __forceinline__ void wave_argmax_dpp(FT& v, int& k) {
    const FT m = wave_reduce_dpp(v, [](FT a, FT b) { return a > b ? a : b; });
    const int c = v == m ? k : INT_MAX;
    k = wave_reduce_dpp(c, [](int a, int b) { return a < b ? a : b; });
    v = m;
}

// Apply the flip of node k: every thread of the calling group takes part
// (stride = group size).  Caller synchronises before and after.
template <typename WT, typename FT>
__device__ __forceinline__ void flip_update(const GView<WT>& g, uint8_t* x, FT* h, int64_t k, int s_old,
                                            int tid, int nthreads) {
    const FT two_s = (FT)(2 * s_old);
    if (g.use_adj) {
        for (int32_t e = g.ap[k] + tid; e < g.ap[k + 1]; e += nthreads) h[g.ai[e]] -= two_s * (FT)g.aw[e];
    } else {
        const WT* Wk = g.W + k * g.ld;
        for (int64_t v = tid; v < g.n; v += nthreads) h[v] -= two_s * (FT)Wk[v];
    }
    if (tid == 0) x[k] ^= 1u;
}

// ---------------------------------------------------------------------------
// Local field h = W s
// ---------------------------------------------------------------------------

template <typename WT, typename FT>
__global__ __launch_bounds__(256) void k_field_dense(GView<WT> g, const uint8_t* x, FT* h) {
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= g.n) return;
    const WT* Wr = g.W + row * g.ld;
    FT acc = 0;
    for (int64_t u = lane; u < g.n; u += 64) acc += x[u] ? -(FT)Wr[u] : (FT)Wr[u];
    acc = wave_sum(acc);
    if (lane == 0) h[row] = acc;
}

template <typename WT, typename FT>
__global__ __launch_bounds__(256) void k_field_adj(GView<WT> g, const uint8_t* x, FT* h) {
    const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= g.n) return;
    FT acc = 0;
    for (int32_t e = g.ap[row]; e < g.ap[row + 1]; ++e) acc += x[g.ai[e]] ? -(FT)g.aw[e] : (FT)g.aw[e];
    h[row] = acc;
}

// the recompute of node `row` replaces h[row]; with a FieldRepair the old
// value is compared first
template <typename FT>
__device__ __forceinline__ void field_store(FT* h, int64_t row, FT exact, FieldRepair* fr) {
    if (fr) {
        const double d = fabs((double)h[row] - (double)exact);
        if (d > 1e-9 * fmax(1.0, fabs((double)exact))) atomicAdd(&fr->entries, 1ull);
        atomicMax(&fr->max_bits, (unsigned long long)__double_as_longlong(d));
    }
    h[row] = exact;
}

template <typename WT, typename FT>
__global__ __launch_bounds__(256) void k_field_repair_dense(GView<WT> g, const uint8_t* x, FT* h, const LsCtl* gate,
                                                            FieldRepair* fr) {
    if (gate && !ls_new_round(gate)) return;
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&fr->checks, 1ull);
    if (row >= g.n) return;
    const WT* Wr = g.W + row * g.ld;
    FT acc = 0;
    for (int64_t u = lane; u < g.n; u += 64) acc += x[u] ? -(FT)Wr[u] : (FT)Wr[u];
    acc = wave_sum(acc);
    if (lane == 0) field_store(h, row, acc, fr);
}

template <typename WT, typename FT>
__global__ __launch_bounds__(256) void k_field_repair_adj(GView<WT> g, const uint8_t* x, FT* h, const LsCtl* gate,
                                                          FieldRepair* fr) {
    if (gate && !ls_new_round(gate)) return;
    const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&fr->checks, 1ull);
    if (row >= g.n) return;
    FT acc = 0;
    for (int32_t e = g.ap[row]; e < g.ap[row + 1]; ++e) acc += x[g.ai[e]] ? -(FT)g.aw[e] : (FT)g.aw[e];
    field_store(h, row, acc, fr);
}

template <typename FT>
__global__ void k_perturb_field(FT* h, int64_t n) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) h[v] += (FT)(1e-3 * (double)((v % 7) - 3));
}

template <typename WT>
static void field_impl(const Graph& g, const uint8_t* x, void* h, hipStream_t s) {
    using FT = typename FieldOf<WT>::T;
    GView<WT> v = view<WT>(g);
    if (g.W && !g.use_adj) {
        hipLaunchKernelGGL((k_field_dense<WT, FT>), dim3((unsigned)((g.n + 3) / 4)), dim3(256), 0, s, v, x, (FT*)h);
    } else {
        hipLaunchKernelGGL((k_field_adj<WT, FT>), dim3((unsigned)((g.n + 255) / 256)), dim3(256), 0, s, v, x, (FT*)h);
    }
}

#define PQI_DISPATCH(g, FN, ...)                                   \
    do {                                                           \
        switch ((g).kind) {                                        \
            case PQI_WEIGHTS_I8: FN<int8_t>(__VA_ARGS__); break;   \
            case PQI_WEIGHTS_F32: FN<float>(__VA_ARGS__); break;   \
            default: FN<double>(__VA_ARGS__); break;               \
        }                                                          \
    } while (0)

// Descent variant: 0 the register-resident packed-key descent where it
// applies, else the full scan; 1 the full scan over keys in LDS; 2 the
// two-level argmax for sparse graphs; 3 the full scan without packed keys for
// integer weights.  Read once from PQI_DESCENT_VARIANT.
static int descent_variant() {
    static const int v = [] {
        const char* e = getenv("PQI_DESCENT_VARIANT");
        return e ? atoi(e) : 0;
    }();
    return v;
}

// Dynamic LDS above 64 KiB must be requested per kernel; gfx950 has 160 KiB
// per work group.  Returns the limit that applies to `fn`.
static size_t lds_limit(const void* fn) {
    const int want = 160 * 1024;
    if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, want) == hipSuccess) return want;
    (void)hipGetLastError();
    return 64 * 1024;
}

void launch_field(const Graph& g, const uint8_t* x, void* h, hipStream_t s) {
    PQI_DISPATCH(g, field_impl, g, x, h, s);
}

template <typename WT>
static void field_repair_impl(const Graph& g, const uint8_t* x, void* h, const LsCtl* gate, FieldRepair* fr,
                              hipStream_t s) {
    using FT = typename FieldOf<WT>::T;
    GView<WT> v = view<WT>(g);
    if (g.W && !g.use_adj) {
        hipLaunchKernelGGL((k_field_repair_dense<WT, FT>), dim3((unsigned)((g.n + 3) / 4)), dim3(256), 0, s, v, x,
                           (FT*)h, gate, fr);
    } else {
        hipLaunchKernelGGL((k_field_repair_adj<WT, FT>), dim3((unsigned)((g.n + 255) / 256)), dim3(256), 0, s, v, x,
                           (FT*)h, gate, fr);
    }
}

void launch_field_repair(const Graph& g, const uint8_t* x, void* h, const LsCtl* gate, FieldRepair* fr,
                         hipStream_t s) {
    if (g.kind == PQI_WEIGHTS_I8) return;  // int32 fields are exact
    PQI_DISPATCH(g, field_repair_impl, g, x, h, gate, fr, s);
}

void launch_perturb_field(const Graph& g, void* h, hipStream_t s) {
    if (g.kind == PQI_WEIGHTS_I8) return;
    hipLaunchKernelGGL(k_perturb_field<double>, dim3((unsigned)((g.n + 255) / 256)), dim3(256), 0, s, (double*)h,
                       g.n);
}

size_t field_bytes(const Graph& g) { return g.kind == PQI_WEIGHTS_I8 ? 4 : 8; }

// ---------------------------------------------------------------------------
// Exact score: q = s^T h  ->  cut = (2 W_up - q)/4, E = -q/2.
// One work group, fixed reduction order (deterministic).
// ---------------------------------------------------------------------------

template <typename FT> struct AccOf { using T = double; };
template <> struct AccOf<int32_t> { using T = int64_t; };

template <typename FT>
__global__ __launch_bounds__(1024) void k_score(const uint8_t* x, const FT* h, int64_t n, double w_up,
                                               int64_t w_up_i, double* out_cut, double* out_e) {
    using AT = typename AccOf<FT>::T;
    __shared__ AT part[16];
    AT acc = 0;
    for (int64_t v = threadIdx.x; v < n; v += blockDim.x) acc += x[v] ? -(AT)h[v] : (AT)h[v];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        AT q = 0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) q += part[i];
        if (sizeof(AT) == 8 && sizeof(FT) == 4) {
            // exact integer arithmetic: 2*W_up - q is a multiple of 4
            int64_t qi = (int64_t)q;
            *out_cut = (double)((2 * w_up_i - qi) / 4);
            *out_e = (double)(-qi / 2);
        } else {
            double qd = (double)q;
            *out_cut = (2.0 * w_up - qd) / 4.0;
            *out_e = -qd / 2.0;
        }
    }
}

template <typename FT>
__device__ void quad_body(const uint8_t* x, const FT* h, int64_t n, double* q) {
    using AT = typename AccOf<FT>::T;
    __shared__ AT part[16];
    AT acc = 0;
    for (int64_t v = threadIdx.x; v < n; v += blockDim.x) acc += x[v] ? -(AT)h[v] : (AT)h[v];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        AT t = 0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += part[i];
        q[0] = (double)t;
    }
}

template <typename FT>
__global__ __launch_bounds__(1024) void k_quad(const uint8_t* x, const FT* h, int64_t n, double* q) {
    quad_body<FT>(x, h, n, q);
}

void launch_quad(const Graph& g, const uint8_t* x, const void* h, double* q, hipStream_t s) {
    if (g.kind == PQI_WEIGHTS_I8)
        hipLaunchKernelGGL((k_quad<int32_t>), dim3(1), dim3(1024), 0, s, x, (const int32_t*)h, g.n, q);
    else
        hipLaunchKernelGGL((k_quad<double>), dim3(1), dim3(1024), 0, s, x, (const double*)h, g.n, q);
}

template <typename WT>
static void score_impl(const Graph& g, const uint8_t* x, int64_t count, double* out_cut, double* out_e,
                       void* hbuf, hipStream_t s) {
    using FT = typename FieldOf<WT>::T;
    for (int64_t b = 0; b < count; ++b) {
        field_impl<WT>(g, x + b * g.n, hbuf, s);
        hipLaunchKernelGGL((k_score<FT>), dim3(1), dim3(1024), 0, s, x + b * g.n, (const FT*)hbuf, g.n, g.w_up,
                           g.w_up_i, out_cut + b, out_e + b);
    }
}

void launch_score_with(const Graph& g, const uint8_t* x, int64_t count, double* out_cut, double* out_e,
                       void* hbuf, hipStream_t s) {
    PQI_DISPATCH(g, score_impl, g, x, count, out_cut, out_e, hbuf, s);
}

// ---------------------------------------------------------------------------
// Single-flip descent: argmax of Delta over all k (first maximum), apply while
// positive.  One work group of 1024 threads.  max_steps = 0 only reports the
// best move (pyqrackising/spin_glass_solver.py:37-52).
// ---------------------------------------------------------------------------

template <typename WT, typename FT>
__device__ void descent_body(const GView<WT>& g, uint8_t* x, FT* h, int64_t max_steps, PhaseOut* out) {
    __shared__ FT sv[16];
    __shared__ int64_t sk[16];
    const int tid = threadIdx.x;
    int64_t steps = 0, evals = 0;
    double gsum = 0.0;
    FT bv;
    int64_t bk;
    for (;;) {
        bv = lowest<FT>();
        bk = LLONG_MAX;
        for (int64_t k = tid; k < g.n; k += 1024) {
            FT d = gain_of(x, h, k);
            if (better(d, k, bv, bk)) { bv = d; bk = k; }
        }
        block_argmax(bv, bk, sv, sk);
        evals += g.n;
        if (!(bv > 0) || steps >= max_steps) break;
        const int s_old = x[bk] ? -1 : 1;
        __syncthreads();
        flip_update(g, x, h, bk, s_old, tid, 1024);
        __syncthreads();
        ++steps;
        gsum += (double)bv;
    }
    if (tid == 0) {
        out->gain = (double)bv;
        out->k = bk;
        out->l = -1;
        out->accepts = steps;
        out->evals = evals;
        out->gain_sum = gsum;
    }
}

template <typename WT, typename FT>
__global__ __launch_bounds__(1024) void k_descent(GView<WT> g, uint8_t* x, FT* h, int64_t max_steps,
                                                 PhaseOut* out, const int32_t* gate) {
    if (gate && *gate) return;
    descent_body<WT, FT>(g, x, h, max_steps, out);
}

// Single-wave staging of global arrays into LDS with 16 loads in flight per
// lane (a plain element loop waits for each load before the next).
// Spins and gains D_v = s_v h_v (and cleared flags when trk is given):
template <typename FT>
__device__ __forceinline__ void stage_gains(const uint8_t* xg, const FT* hg, int n, uint8_t* x, FT* D,
                                            uint8_t* trk) {
    const int lane = threadIdx.x;
    for (int v0 = lane; v0 < n; v0 += 64 * 16) {
        uint8_t xv[16];
        FT hv[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int v = v0 + 64 * u;
            xv[u] = v < n ? xg[v] : 0;
            hv[u] = v < n ? hg[v] : (FT)0;
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int v = v0 + 64 * u;
            if (v < n) {
                x[v] = xv[u];
                D[v] = xv[u] ? -hv[u] : hv[u];
                if (trk) trk[v] = 0;
            }
        }
    }
}

// cnt 32-bit words
__device__ __forceinline__ void stage_words(const int32_t* src, int cnt, int32_t* dst) {
    const int lane = threadIdx.x;
    for (int i0 = lane; i0 < cnt; i0 += 64 * 16) {
        int32_t w[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) w[u] = i0 + 64 * u < cnt ? src[i0 + 64 * u] : 0;
#pragma unroll
        for (int u = 0; u < 16; ++u)
            if (i0 + 64 * u < cnt) dst[i0 + 64 * u] = w[u];
    }
}

// Single-wavefront variant.  LDS holds the gains D_v = s_v h_v themselves, so
// the argmax scan is one independent LDS load per element (unrolled), and a
// flip of k updates D_k -> -D_k and D_v -= 2 w_vk s_k s_v for its neighbours.
// No work-group barriers, 32-bit indices; same selection rule and result.
template <typename WT, typename FT>
__device__ void descent_wave_body(const GView<WT>& g, uint8_t* xg, FT* hg, int64_t max_steps, PhaseOut* out,
                                  unsigned char* lds, int stage_ap = 0) {
    const int lane = threadIdx.x;
    const int n = (int)g.n;
    FT* D = (FT*)lds;
    uint8_t* x = lds + (((size_t)n * sizeof(FT) + 15) & ~(size_t)15);
    // adjacency row pointers staged after x (stage_ap: the caller sized LDS
    // for them), so a flip waits for one global access (its row), not two
    int32_t* AP = (int32_t*)(lds + (((size_t)n * sizeof(FT) + 15) & ~(size_t)15) + (((size_t)n + 15) & ~(size_t)15));
    const int32_t* ap = stage_ap ? AP : g.ap;
    stage_gains<FT>(xg, hg, n, x, D, nullptr);
    if (stage_ap) stage_words(g.ap, n + 1, AP);
    __syncthreads();
    int64_t steps = 0, evals = 0;
    double gsum = 0.0;
    FT bv;
    int bk;
    const int full = n & ~511;  // unrolled part: 8 elements per lane per pass
    for (;;) {
        bv = lowest<FT>();
        bk = INT_MAX;
        for (int v0 = lane; v0 < full; v0 += 512) {
            FT d[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) d[u] = D[v0 + 64 * u];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (d[u] > bv) { bv = d[u]; bk = v0 + 64 * u; }
        }
        for (int v = full + lane; v < n; v += 64) {
            const FT d = D[v];
            if (d > bv) { bv = d; bk = v; }
        }
        wave_argmax_dpp(bv, bk);
        evals += n;
        if (!(bv > 0) || steps >= max_steps) break;
        const int sk = x[bk] ? -1 : 1;  // s_k before the flip
        __syncthreads();
        if (g.use_adj) {
            for (int32_t e = ap[bk] + lane; e < ap[bk + 1]; e += 64) {
                const int v = g.ai[e];
                const int sv = x[v] ? -1 : 1;
                D[v] -= (FT)(2 * sk * sv) * (FT)g.aw[e];
            }
        } else {
            const WT* Wk = g.W + (int64_t)bk * g.ld;
            for (int v = lane; v < n; v += 64) {
                if (v == bk) continue;
                const int sv = x[v] ? -1 : 1;
                D[v] -= (FT)(2 * sk * sv) * (FT)Wk[v];
            }
        }
        __syncthreads();
        if (lane == 0) {
            D[bk] = -D[bk];
            x[bk] ^= 1u;
        }
        __syncthreads();
        ++steps;
        gsum += (double)bv;
    }
    if (steps) {
        for (int v = lane; v < n; v += 64) {
            const uint8_t xv = x[v];
            hg[v] = xv ? -D[v] : D[v];
            xg[v] = xv;
        }
    }
    if (lane == 0) {
        out->gain = (double)bv;
        out->k = bk;
        out->l = -1;
        out->accepts = steps;
        out->evals = evals;
        out->gain_sum = gsum;
    }
}

// Sparse-graph descent in rounds: the same moves as descent_wave_body with
// a full scan only once per round.  The scan gives each lane the 4th largest
// gain among its nodes; T, the largest of these over the lanes, bounds every
// node outside the round's candidates {v : D_v > T} (at most 3 per lane).
// Their neighbour rows are fetched into registers in three batches of loads.
// A flip that raises an untracked neighbour's gain above T adds it to the "extras",
// so every untracked node stays <= T, and while the best tracked gain is
// above T its first maximum (lowest node among equal gains) is the first
// maximum over all n nodes: the move of the full scan
// (pyqrackising/spin_glass_solver.py:37-52).  The round ends when it is not.
// When the top gain itself does not exceed T (ties), the scan's own first
// maximum is flipped and a new round starts.  Single wave: LDS accesses of
// one wave complete in order, so phases are separated by compiler barriers
// only.  A flip reads its row from registers (global memory only for an
// extra's row, or rows longer than RD_RS).
constexpr int RD_C = 3;    // candidates per lane (nodes above a lane's 4th largest gain)
constexpr int RD_RS = 16;  // cached entries per candidate row
constexpr int RD_X = 64;   // extras per round

struct RoundLds {
    size_t d, x, trk, ap, sn, ex, total;
};

__device__ __forceinline__ double readlane_t(double v, int l) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                            __builtin_amdgcn_readlane(__double2loint(v), l));
}
__device__ __forceinline__ int32_t readlane_t(int32_t v, int l) { return __builtin_amdgcn_readlane(v, l); }

// maximum / minimum as one instruction (v_max_f64 etc.); gains are never NaN
__device__ __forceinline__ double vmax(double a, double b) { return __builtin_fmax(a, b); }
__device__ __forceinline__ double vmin(double a, double b) { return __builtin_fmin(a, b); }
__device__ __forceinline__ int32_t vmax(int32_t a, int32_t b) { return max(a, b); }
__device__ __forceinline__ int32_t vmin(int32_t a, int32_t b) { return min(a, b); }

// An order-preserving int32 key of a gain: the high word of a double with
// its magnitude bits flipped when negative (so signed integer order follows
// the value), or an integer gain itself.  Keys are coarser than gains, never
// out of order.
__device__ __forceinline__ int32_t gain_key(const double* D, int v) {
    const int32_t hi = ((const int32_t*)D)[2 * v + 1];
    return hi ^ ((hi >> 31) & 0x7fffffff);
}
__device__ __forceinline__ int32_t gain_key(const int32_t* D, int v) { return D[v]; }
__device__ __forceinline__ int32_t hi_key(int32_t hi) { return hi ^ ((hi >> 31) & 0x7fffffff); }

// the key of one gain held in registers
__device__ __forceinline__ int32_t gain_key_of(double d) { return hi_key(__double2hiint(d)); }
__device__ __forceinline__ int32_t gain_key_of(int32_t d) { return d; }

// keys of the gains of nodes b and b + 1 (b even; one 16- or 8-byte load)
__device__ __forceinline__ void gain_key_pair(const double* D, int b, int32_t& k0, int32_t& k1) {
    const int4 q = *(const int4*)(D + b);
    k0 = hi_key(q.y);
    k1 = hi_key(q.w);
}
__device__ __forceinline__ void gain_key_pair(const int32_t* D, int b, int32_t& k0, int32_t& k1) {
    const int2 q = *(const int2*)(D + b);
    k0 = q.x;
    k1 = q.y;
}

// The largest gain whose key is at most k (k = INT_MIN: below every gain)
__device__ __forceinline__ void key_bound(int32_t k, double& t) {
    if (k == INT_MIN) {
        t = -INFINITY;
        return;
    }
    t = k < 0 ? __hiloint2double(k ^ 0x7fffffff, 0) : __hiloint2double(k, (int)0xFFFFFFFF);
}
__device__ __forceinline__ void key_bound(int32_t k, int32_t& t) { t = k; }

// The round descent's per-flip chain has two LDS round trips (the tracked
// gains; the neighbours' spins, flags and gains).  The round's candidates'
// neighbour rows live in registers: candidate c = lane +
// 64 r (r < RD_C) has entry j of its row in register c >> 2 (banks of 24) of
// lane 16 (c & 3) + j, so the winner's row is one uniform register read and
// the 16 lanes of its group update one neighbour each.  The spins of a lane's
// tracked nodes are bits of a register.  A flip: re-read the tracked gains
// (LDS), each lane's exact best (highest gain, then lowest node), a wave
// maximum of the best gains' int32 keys, the winner by ballot (exact
// comparisons only when lanes share the top key), then the neighbours'
// spins, flags and gains (LDS) and their new gains.  Scans and row fetches
// load without bounds tests (the gains are padded to a multiple of 512 nodes
// with the lowest gain), so every load of a stage is in flight at once.
constexpr int RR_NQ = 64 * RD_C / 4;  // row registers per lane (48)
constexpr int RR_BANK = RR_NQ / 2;    // banks of 24: a uniform index into each stays a register move

// a weight as 32-bit words and back (8-byte weights take two)
__device__ __forceinline__ void weight_words(float w, uint32_t& w0, uint32_t& w1) { w0 = __float_as_uint(w); }
__device__ __forceinline__ void weight_words(double w, uint32_t& w0, uint32_t& w1) {
    w0 = (uint32_t)__double2loint(w);
    w1 = (uint32_t)__double2hiint(w);
}
__device__ __forceinline__ void weight_words(int8_t w, uint32_t& w0, uint32_t& w1) { w0 = (uint32_t)(int32_t)w; }
template <typename WT>
__device__ __forceinline__ WT weight_of_words(uint32_t w0, uint32_t w1);
template <>
__device__ __forceinline__ float weight_of_words<float>(uint32_t w0, uint32_t) { return __uint_as_float(w0); }
template <>
__device__ __forceinline__ double weight_of_words<double>(uint32_t w0, uint32_t w1) {
    return __hiloint2double((int)w1, (int)w0);
}
template <>
__device__ __forceinline__ int8_t weight_of_words<int8_t>(uint32_t w0, uint32_t) { return (int8_t)(int32_t)w0; }

template <typename WT, typename FT>
__host__ __device__ inline RoundLds round_lds_layout(int64_t n) {
    auto al = [](size_t v) { return (v + 15) & ~(size_t)15; };
    RoundLds L;
    size_t o = 0;
    L.d = o;
    // + a sentinel below every gain at n, and the same up to a multiple of
    // 512 nodes so that the scans load without bounds tests
    o += al((size_t)((n + 1 + 511) & ~(int64_t)511) * sizeof(FT));
    L.x = o;
    o += al((size_t)n);
    L.trk = o;
    o += al((size_t)n);
    L.ap = o;
    o += al((size_t)(n + 1) * 4);
    L.sn = o;
    o += al(64 * RD_C * 4);
    L.ex = o;
    o += al(RD_X * 4);
    L.total = o;
    return L;
}

template <typename WT, typename FT>
__device__ void descent_round_body(const GView<WT>& g, uint8_t* xg, FT* hg, int64_t max_steps, PhaseOut* out,
                                    unsigned char* lds) {
    constexpr int RS = RD_RS;
    static_assert(RS == 16 && RD_C == 3, "16 lanes per row, 4 rows per register");
    const int lane = threadIdx.x;
    const int n = (int)g.n;
    const RoundLds L = round_lds_layout<WT, FT>(n);
    FT* D = (FT*)(lds + L.d);
    uint8_t* x = lds + L.x;      // the spins (1: -1)
    uint8_t* trk = lds + L.trk;  // 1: tracked in this round
    int32_t* AP = (int32_t*)(lds + L.ap);
    int32_t* SN = (int32_t*)(lds + L.sn);  // the round's candidates
    int32_t* EX = (int32_t*)(lds + L.ex);
    auto fmax_ = [](FT a, FT b) { return vmax(a, b); };
    auto imin_ = [](int a, int b) { return a < b ? a : b; };
    stage_gains<FT>(xg, hg, n, x, D, trk);
    stage_words(g.ap, n + 1, AP);
    const int nscan = (n + 511) >> 9;  // scan iterations of 512 nodes
    for (int v = n + lane; v < ((n + 1 + 511) & ~511); v += 64) D[v] = lowest<FT>();  // sentinel and padding
    __syncthreads();
    // 32-bit step counters (a descent makes far fewer than 2^31 moves)
    const int32_t step_cap = max_steps < INT_MAX ? (int32_t)max_steps : INT_MAX;
    int32_t steps = 0;
    double gsum = 0.0;
    FT bv = lowest<FT>();
    int bk = INT_MAX;
    FT T = lowest<FT>();
    int nx = 0;  // extras this round (wave-uniform)
    // the gain update of one neighbour per active lane (v, w) for a flip with
    // spin s_k; with tracking, a neighbour raised above T joins the extras.
    // Returns false when the extras overflowed (the round must end).
    auto update = [&](bool act, int v, WT w, int sk, bool tracking) -> bool {
        bool add = false;
        if (act) {
            const uint8_t xv = x[v], tv = trk[v];
            const FT nd = D[v] - (FT)(2 * sk * (xv ? -1 : 1)) * (FT)w;
            D[v] = nd;
            add = tracking && !tv && nd > T;
        }
        const uint64_t m = __ballot(add);
        if (m) {
            const int cnt = __popcll(m);
            if (nx + cnt > RD_X) return false;
            if (add) {
                const int pos = nx + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                   __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                EX[pos] = v;
                trk[v] = 1;
            }
            nx += cnt;
        }
        return true;
    };
    // a flip whose row comes from global memory, lanes over the entries
    auto flip_global = [&](int k, int sk, bool tracking) -> bool {
        const int e0 = AP[k], deg = AP[k + 1] - e0;
        bool fits = true;
        for (int j0 = 0; j0 < deg; j0 += 64) {
            const int j = j0 + lane;
            const bool act = j < deg;
            const int v = act ? g.ai[e0 + j] : 0;
            const WT w = act ? g.aw[e0 + j] : (WT)0;
            fits &= update(act, v, w, sk, tracking);
        }
        return fits;
    };
    // every lane stores the same values at the same addresses (no branch)
    auto set_k = [&](int k, FT gk, int xk) {
        __builtin_amdgcn_wave_barrier();
        D[k] = -gk;
        x[k] = (uint8_t)(xk ^ 1);
        __builtin_amdgcn_wave_barrier();
    };
    // the round's candidate rows (banks of 24: a uniform index into each is a
    // register move)
    // weights as 32-bit words (8-byte weights in two banks of words)
    constexpr int WW = sizeof(WT) == 8 ? 2 : 1;
    int32_t riA[RR_BANK], riB[RR_BANK];
    uint32_t wA0[RR_BANK], wB0[RR_BANK], wA1[WW == 2 ? RR_BANK : 1], wB1[WW == 2 ? RR_BANK : 1];
    for (;;) {
        // scan: the lane's 4 largest gain keys (a max/min insertion network);
        // KT is the largest 4th over the lanes and T the largest gain with a
        // key <= KT, so a node outside {key > KT} has a gain <= T, and each
        // lane has at most 3 nodes with key > KT
        int32_t c0 = INT_MIN, c1 = INT_MIN, c2 = INT_MIN, c3 = INT_MIN;
        auto insert = [&](int32_t k) {
            const int32_t m0 = min(c0, k);
            c0 = max(c0, k);
            const int32_t m1 = min(c1, m0);
            c1 = max(c1, m0);
            const int32_t m2 = min(c2, m1);
            c2 = max(c2, m1);
            c3 = max(c3, m2);
        };
        // lane l holds the pairs of nodes 2 l + 128 i and 2 l + 128 i + 1
        // (the padding past n holds the lowest gain: loads without tests)
        for (int it = 0; it < nscan; ++it) {
            const int b0 = 2 * lane + 512 * it;
            int32_t ka[4], kb[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) gain_key_pair(D, b0 + 128 * u, ka[u], kb[u]);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                insert(ka[u]);
                insert(kb[u]);
            }
        }
        const int32_t K0 = wave_max_i32(c0);
        const int32_t KT = wave_max_i32(c3);
        key_bound(KT, T);
        if (K0 <= 0 || K0 <= KT || steps >= step_cap) {
            // the first maximum itself (exact): the result, or the move to
            // make directly when no key exceeds KT (ties) or gains are tiny
            FT lb = lowest<FT>();
            int lk = INT_MAX;
            for (int v = lane; v < n; v += 64) {
                const FT d = D[v];
                if (d > lb) {
                    lb = d;
                    lk = v;
                }
            }
            wave_argmax_dpp(lb, lk);
            bv = lb;
            bk = lk;
            if (lk == INT_MAX || !(lb > 0) || steps >= step_cap) break;
            const int xk = x[lk];
            flip_global(lk, xk ? -1 : 1, false);
            set_k(lk, lb, xk);
            ++steps;
            gsum += (double)lb;
            continue;
        }
        // the candidates: key > KT, at most RD_C per lane (same lane layout
        // as the scan), in ascending node order; their spins as bits
        int cn[RD_C], ncl = 0;
#pragma unroll
        for (int r = 0; r < RD_C; ++r) cn[r] = -1;
        auto take = [&](bool ok, int v) {
#pragma unroll
            for (int r = 0; r < RD_C; ++r) cn[r] = (ok && ncl == r) ? v : cn[r];
            ncl += ok ? 1 : 0;
        };
        for (int it = 0; it < nscan; ++it) {
            const int b0 = 2 * lane + 512 * it;
            int32_t ka[4], kb[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) gain_key_pair(D, b0 + 128 * u, ka[u], kb[u]);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                take(ka[u] > KT, b0 + 128 * u);
                take(kb[u] > KT, b0 + 128 * u + 1);
            }
        }
        uint32_t xs = 0;  // bit r: spin of slot r (RD_C: the extra)
        int cdg[RD_C];
#pragma unroll
        for (int r = 0; r < RD_C; ++r) {
            SN[lane + 64 * r] = cn[r];
            cdg[r] = 0;
            if (cn[r] >= 0) {
                trk[cn[r]] = 1;
                xs |= (uint32_t)(x[cn[r]] != 0) << r;
                cdg[r] = AP[cn[r] + 1] - AP[cn[r]];
            }
        }
        __builtin_amdgcn_wave_barrier();
        // the candidates' rows into registers: register q of lane 16 g + j
        // holds entry j of candidate 4 q + g (node, or -1 past the row)
#pragma unroll
        for (int q0 = 0; q0 < RR_NQ; q0 += 16) {
            int32_t e0[16], dg[16];
            // reads without tests (an empty slot reads row 0 and keeps no
            // entry), so every read of a stage is in flight at once
            int32_t cs[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) cs[u] = SN[4 * (q0 + u) + (lane >> 4)];
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int c = cs[u] >= 0 ? cs[u] : 0;
                e0[u] = AP[c];
                dg[u] = AP[c + 1];
            }
#pragma unroll
            for (int u = 0; u < 16; ++u) dg[u] = cs[u] >= 0 ? dg[u] - e0[u] : 0;
            int32_t vi[16];
            WT vw[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                // loads without tests (entry 0 past a row), so all are in flight
                const bool ok = (lane & 15) < dg[u];
                const int e = ok ? e0[u] + (lane & 15) : 0;
                vi[u] = g.ai[e];
                vw[u] = g.aw[e];
            }
#pragma unroll
            for (int u = 0; u < 16; ++u) vi[u] = (lane & 15) < dg[u] ? vi[u] : -1;
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int q = q0 + u;
                uint32_t w0, w1 = 0;
                weight_words(vw[u], w0, w1);
                if (q < RR_BANK) {
                    riA[q < RR_BANK ? q : 0] = vi[u];
                    wA0[q < RR_BANK ? q : 0] = w0;
                    if constexpr (WW == 2) wA1[q < RR_BANK ? q : 0] = w1;
                } else {
                    riB[q >= RR_BANK ? q - RR_BANK : 0] = vi[u];
                    wB0[q >= RR_BANK ? q - RR_BANK : 0] = w0;
                    if constexpr (WW == 2) wB1[q >= RR_BANK ? q - RR_BANK : 0] = w1;
                }
            }
        }
        nx = 0;
        bool term = false;  // a termination condition met: the next scan settles it exactly
        int32_t round_flips = 0;
        int myex = -1;  // this lane's extra (EX[lane])
        // the slots' gain addresses (the sentinel at n for empty slots), kept
        // ready so that the four reads of a step go out together
        int cnn[RD_C];
#pragma unroll
        for (int r = 0; r < RD_C; ++r) cnn[r] = cn[r] >= 0 ? cn[r] : n;
        int exn = n;
        for (;;) {
            // the tracked gains (empty slots read the sentinel), all in flight
            FT d[RD_C + 1];
#pragma unroll
            for (int r = 0; r < RD_C; ++r) d[r] = D[cnn[r]];
            d[RD_C] = D[exn];
            // the wave's first maximum: a maximum of the gain keys, and when
            // exactly one tracked node holds that key it is the winner;
            // otherwise exact comparisons among the nodes holding it
            int32_t kr[RD_C + 1];
#pragma unroll
            for (int r = 0; r <= RD_C; ++r) kr[r] = gain_key_of(d[r]);
            const int32_t K = wave_max_i32(max(max(kr[0], kr[1]), max(kr[2], kr[3])));
            // each lane: how many of its slots hold K, and the first; one
            // ballot finds the lanes, and a single holder is the winner
            const int h0 = kr[0] == K, h1 = kr[1] == K, h2 = kr[2] == K, h3 = kr[3] == K;
            const int lane_hits = h0 + h1 + h2 + h3;
            const int first = h0 ? 0 : (h1 ? 1 : (h2 ? 2 : 3));
            const uint64_t holders = __ballot(lane_hits != 0);
            int w = (int)__builtin_ctzll(holders);
            const int wp = __builtin_amdgcn_readlane(first | (lane_hits << 2), w);
            int rw = wp & 3;
            if (((wp >> 2) != 1) | (__popcll(holders) != 1)) {
                // several nodes share the top key: the largest gain among them,
                // then the lowest node
                FT bd = lowest<FT>();
                int bn = INT_MAX, br = 0;
#pragma unroll
                for (int r = 0; r <= RD_C; ++r) {
                    const int nd = r < RD_C ? cn[r] : myex;
                    if (kr[r] == K && (d[r] > bd || (d[r] == bd && nd < bn))) {
                        bd = d[r];
                        bn = nd;
                        br = r;
                    }
                }
                const FT mm = wave_reduce_dpp(bd, fmax_);
                const int c = bd == mm ? bn : INT_MAX;
                const int mn = wave_reduce_dpp(c, imin_);
                w = (int)__builtin_ctzll(__ballot(bn == mn && bd == mm));
                rw = __builtin_amdgcn_readlane(br, w);
            }
            // the winner's gain, node and spin (rw is wave-uniform)
            const FT dsel = rw == 0 ? d[0] : (rw == 1 ? d[1] : (rw == 2 ? d[2] : d[3]));
            const int nsel = rw == 0 ? cn[0] : (rw == 1 ? cn[1] : (rw == 2 ? cn[2] : myex));
            const FT m = readlane_t(dsel, w);
            const int k = __builtin_amdgcn_readlane(nsel, w);
            const int xk = (int)((__builtin_amdgcn_readlane(xs, w) >> rw) & 1u);
            if (!(m > T)) break;  // an untracked node may hold the maximum: rescan
            if (!(m > 0) || steps >= step_cap) {
                term = true;
                break;
            }
            const int nx0 = nx;
            bool fits;
            int deg = RS + 1;
            if (rw < RD_C) {
                const int cg = rw == 0 ? cdg[0] : (rw == 1 ? cdg[1] : cdg[2]);
                deg = __builtin_amdgcn_readlane(cg, w);
            }
            if (deg <= RS) {
                // the row from registers: candidate c, one entry per lane of
                // its group; the other lanes read and write back the sentinel
                const int c = w + 64 * rw;
                const int q = c >> 2;
                // one read from each bank, then a select of the values (a
                // select of the two arrays' addresses would put them in memory)
                const int qa = q < RR_BANK ? q : 0, qb = q >= RR_BANK ? q - RR_BANK : 0;
                int32_t ia = riA[qa], ib = riB[qb];
                uint32_t a0 = wA0[qa], b0 = wB0[qb], a1 = 0, b1 = 0;
                if constexpr (WW == 2) {
                    a1 = wA1[qa];
                    b1 = wB1[qb];
                }
                __asm__("" : "+v"(ia), "+v"(ib), "+v"(a0), "+v"(b0), "+v"(a1), "+v"(b1));
                const int vi = q < RR_BANK ? ia : ib;
                const FT w2 = (FT)2 * (FT)weight_of_words<WT>(q < RR_BANK ? a0 : b0, q < RR_BANK ? a1 : b1);
                const bool act = (lane >> 4) == (c & 3) && vi >= 0;
                const int v = act ? vi : n;
                const uint8_t xv = x[v], tv = trk[v];
                const FT dv = D[v];
                // D_v - 2 w s_k s_v (s_k s_v = 1 when the spins are equal);
                // the sentinel stays the lowest gain (-inf, or kept as is)
                const FT t = (int)(xv != 0) == xk ? w2 : -w2;
                FT nd = dv - t;
                if constexpr (!std::is_floating_point<FT>::value) nd = act ? nd : dv;
                D[v] = nd;
                const bool add = act && !tv && nd > T;
                const uint64_t am = __ballot(add);
                fits = true;
                if (am) {
                    const int cnt = __popcll(am);
                    if (nx + cnt > RD_X) {
                        fits = false;
                    } else {
                        if (add) {
                            const int pos = nx + (int)__builtin_amdgcn_mbcnt_hi(
                                                     (uint32_t)(am >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)am, 0u));
                            EX[pos] = v;
                            trk[v] = 1;
                        }
                        nx += cnt;
                    }
                }
            } else {
                fits = flip_global(k, xk ? -1 : 1, true);
            }
            set_k(k, m, xk);
            if (lane == w) xs ^= 1u << rw;
            if (nx > nx0 && lane >= nx0 && lane < nx) {
                myex = EX[lane];
                exn = myex;
                xs = (xs & ~(1u << RD_C)) | ((uint32_t)(x[myex] != 0) << RD_C);
            }
            __builtin_amdgcn_wave_barrier();
            ++steps;
            ++round_flips;
            gsum += (double)m;
            if (!fits) break;
        }
        // untrack the round's candidates and extras
#pragma unroll
        for (int r = 0; r < RD_C; ++r)
            if (cn[r] >= 0) trk[cn[r]] = 0;
        if (myex >= 0) trk[myex] = 0;
        __builtin_amdgcn_wave_barrier();
        if (round_flips == 0 && !term) {
            // cannot happen (the node of key K0 > KT is a candidate above T);
            // flipping the first maximum keeps every round making progress
            FT lb = lowest<FT>();
            int lk = INT_MAX;
            for (int v = lane; v < n; v += 64) {
                const FT d = D[v];
                if (d > lb) {
                    lb = d;
                    lk = v;
                }
            }
            wave_argmax_dpp(lb, lk);
            bv = lb;
            bk = lk;
            if (lk == INT_MAX || !(lb > 0)) break;
            const int xk = x[lk];
            flip_global(lk, xk ? -1 : 1, false);
            set_k(lk, lb, xk);
            ++steps;
            gsum += (double)lb;
        }
    }
    __syncthreads();
    if (steps) {
        for (int v = lane; v < n; v += 64) {
            const uint8_t xv = x[v];
            hg[v] = xv ? -D[v] : D[v];
            xg[v] = xv;
        }
    }
    if (lane == 0) {
        out->gain = (double)bv;
        out->k = bk;
        out->l = -1;
        out->accepts = steps;
        out->evals = (int64_t)n * (steps + 1);
        out->gain_sum = gsum;
    }
}

// Integer-weight descent on packed keys.
//   key_v = (Delta_v << (ib + 1)) | ((mask - v) << 1) | x_v,  mask = 2^ib - 1 >= n - 1.
// The largest key carries the largest gain and, among equal gains, the lowest
// node -- the first argmax of pyqrackising/spin_glass_solver.py:45 as one
// 32-bit integer maximum (v_max3 over 4-wide LDS loads, then a DPP
// reduction) -- and the spin rides in bit 0, so a flip is one read-modify-
// write per neighbour.  The host enables this only when every row strength
// sum_u |w_vu| is below 2^(30 - ib), so no key overflows at any point of the
// descent.  Neighbours come from LDS when they fit there: a fixed-stride table
// (row v at v * stride, entries (w << 24) | v', zero padded) built at upload,
// or the CSR lists; otherwise from global adjacency lists or dense rows.
// Same selection rule and result as the other descents.
struct KeyLds {
    int np;          // n rounded up to 256: key slots (padding holds INT_MIN)
    size_t base;     // bytes of the keys alone
    size_t fixed;    // + fixed-stride table (0 if none)
    size_t csr;      // + CSR row pointers and entries (0 if no lists)
};

__host__ __device__ inline KeyLds key_lds_layout(int64_t n, int64_t nnz, int64_t stride, bool lists) {
    KeyLds L;
    L.np = (int)((n + 255) & ~(int64_t)255);
    L.base = (size_t)L.np * 4;
    L.fixed = stride > 0 ? L.base + (((size_t)n * (size_t)stride + 3) & ~(size_t)3) * 4 : 0;
    L.csr = lists ? L.base + (((size_t)n + 1 + 3) & ~(size_t)3) * 4 + (size_t)nnz * 4 : 0;
    return L;
}

// 0: unstaged, 1: fixed-stride table in LDS, 2: CSR lists in LDS
__host__ __device__ inline int key_stage_mode(const KeyLds& L, size_t lds_bytes) {
    if (L.fixed && L.fixed <= lds_bytes) return 1;
    if (L.csr && L.csr <= lds_bytes) return 2;
    return 0;
}

__device__ __forceinline__ int32_t key_of(int32_t d, int v, uint32_t x, int ib, int mask) {
    return (int32_t)(((uint32_t)d << (ib + 1)) | ((uint32_t)(mask - v) << 1) | x);
}

// K[v] -= 2 s_k s_v w << (ib + 1), with s_v read from the key itself
__device__ __forceinline__ void key_update(int32_t* K, int v, int sk2w, int ib) {
    const int32_t k = K[v];
    K[v] = (int32_t)((uint32_t)k - ((uint32_t)((k & 1) ? -sk2w : sk2w) << (ib + 1)));
}

__device__ void descent_key_body(const GView<int8_t>& g, uint8_t* xg, int32_t* hg, int64_t max_steps,
                                 PhaseOut* out, unsigned char* lds, size_t lds_bytes, int ib) {
    const int lane = threadIdx.x;
    const int n = (int)g.n;
    const int mask = (1 << ib) - 1;
    const int sh = ib + 1;
    const KeyLds L = key_lds_layout(n, g.use_adj ? g.ap[n] : 0, g.kstride, g.use_adj != 0);
    const int mode = max_steps > 0 ? key_stage_mode(L, lds_bytes) : 0;
    int32_t* K = (int32_t*)lds;
    int32_t* T = (int32_t*)(lds + L.base);  // mode 1: fixed table; mode 2: row pointers
    int32_t* A = T + ((n + 1 + 3) & ~3);    // mode 2: entries
    for (int v = lane; v < L.np; v += 64) {
        if (v < n) {
            const uint8_t xv = xg[v];
            K[v] = key_of(xv ? -hg[v] : hg[v], v, xv, ib, mask);
        } else {
            K[v] = INT_MIN;
        }
    }
    if (mode == 1) {
        const int quads = (n * (int)g.kstride + 3) >> 2;  // the table is zero padded to whole quads
        const int4* src = (const int4*)g.kadj;
        int4* dst = (int4*)T;
#pragma unroll 12  // 12 loads in flight per lane: a 24 KiB table in two rounds
        for (int q = lane; q < quads; q += 64) dst[q] = src[q];
    } else if (mode == 2) {
        for (int v = lane; v <= n; v += 64) T[v] = g.ap[v];
        const int nnz = g.ap[n];
        for (int e = lane; e < nnz; e += 64) A[e] = (int32_t)((uint32_t)(uint8_t)g.aw[e] << 24) | g.ai[e];
    }
    __syncthreads();
    int64_t steps = 0, evals = 0;
    double gsum = 0.0;
    int32_t bv;
    int bk;
    const int4* K4 = (const int4*)K;
    const int chunks = L.np >> 2;  // multiple of 64
    const int stride = (int)g.kstride;
    for (;;) {
        // eight independent 16-byte loads in flight per lane
        int32_t m0 = INT_MIN, m1 = INT_MIN, m2 = INT_MIN, m3 = INT_MIN;
        int c = lane;
        for (; c + 448 < chunks; c += 512) {
            int4 a[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) a[u] = K4[c + 64 * u];
            m0 = max(m0, max(max3i(a[0].x, a[0].y, a[0].z), max3i(a[0].w, a[1].x, a[1].y)));
            m1 = max(m1, max(max3i(a[1].z, a[1].w, a[2].x), max3i(a[2].y, a[2].z, a[2].w)));
            m2 = max(m2, max(max3i(a[3].x, a[3].y, a[3].z), max3i(a[3].w, a[4].x, a[4].y)));
            m3 = max(m3, max(max3i(a[4].z, a[4].w, a[5].x), max3i(a[5].y, a[5].z, a[5].w)));
            m0 = max(m0, max(max3i(a[6].x, a[6].y, a[6].z), max3i(a[6].w, a[7].x, a[7].y)));
            m1 = max(m1, max(a[7].z, a[7].w));
        }
        for (; c < chunks; c += 64) {
            const int4 a = K4[c];
            m2 = max(m2, max(max(a.x, a.y), max(a.z, a.w)));
        }
        const int32_t m = wave_max_i32(max(max(m0, m1), max(m2, m3)));
        bv = m >> sh;
        bk = mask - ((m >> 1) & mask);
        evals += n;
        if (!(bv > 0) || steps >= max_steps) break;
        const uint32_t xk = (uint32_t)m & 1u;
        const int sk2 = xk ? -2 : 2;  // 2 s_k before the flip
        if (mode == 1) {
            for (int j = lane; j < stride; j += 64) {
                const int32_t a = T[bk * stride + j];
                if (a != 0) key_update(K, a & 0xFFFFFF, sk2 * (a >> 24), ib);
            }
        } else if (mode == 2) {
            const int beg = T[bk], end = T[bk + 1];
            for (int e = beg + lane; e < end; e += 64) {
                const int32_t a = A[e];
                key_update(K, a & 0xFFFFFF, sk2 * (a >> 24), ib);
            }
        } else if (g.use_adj) {
            const int beg = g.ap[bk], end = g.ap[bk + 1];
            for (int e = beg + lane; e < end; e += 64) key_update(K, g.ai[e], sk2 * (int)g.aw[e], ib);
        } else {
            // dense row, 4 nodes per lane per pass: W is zero beyond n and on
            // the diagonal, and the np >= n_pad key slots cover the row
            const char4* W4 = (const char4*)(g.W + (int64_t)bk * g.ld);
            for (int q = lane; q < (int)(g.ld >> 2); q += 64) {
                const char4 w = W4[q];
                const int v = q * 4;
                if (w.x) key_update(K, v, sk2 * w.x, ib);
                if (w.y) key_update(K, v + 1, sk2 * w.y, ib);
                if (w.z) key_update(K, v + 2, sk2 * w.z, ib);
                if (w.w) key_update(K, v + 3, sk2 * w.w, ib);
            }
        }
        if (lane == 0) K[bk] = key_of(-bv, bk, xk ^ 1u, ib, mask);
        __syncthreads();
        ++steps;
        gsum += (double)bv;
    }
    if (steps) {
        for (int v = lane; v < n; v += 64) {
            const int32_t k = K[v];
            const int32_t d = k >> sh;
            hg[v] = (k & 1) ? -d : d;
            xg[v] = (uint8_t)(k & 1);
        }
    }
    if (lane == 0) {
        out->gain = (double)bv;
        out->k = bk;
        out->l = -1;
        out->accepts = steps;
        out->evals = evals;
        out->gain_sum = gsum;
    }
}

__global__ __launch_bounds__(64) void k_descent_key(GView<int8_t> g, uint8_t* xg, int32_t* hg, int64_t max_steps,
                                                   PhaseOut* out, size_t lds_bytes, int ib, const int32_t* gate) {
    extern __shared__ __align__(16) unsigned char lds[];
    if (gate && *gate) return;
    descent_key_body(g, xg, hg, max_steps, out, lds, lds_bytes, ib);
}

// Register-resident form of the packed-key descent for graphs with a
// fixed-stride neighbour table of at most REG_NB entries per row.  Node v's key
// lives in register v >> 6 of lane v & 63 (n <= 64 R), so a step's maximum is
// a v_max3 tree over R registers and one DPP reduction with no LDS access on
// the critical path.  While the reduction runs, every lane already loads the
// neighbour row of its own local maximum; the winning lane (bk & 63) then
// holds the row of bk in registers.  A flip updates the neighbours one at a
// time: the register index u >> 6 is uniform (v_movrel), and only lane u & 63
// writes.  Same keys, selection rule, counters and result as descent_key_body.
constexpr int REG_NB = 16;

template <int R, int NB>
__device__ void descent_reg_body(const GView<int8_t>& g, uint8_t* xg, int32_t* hg, int64_t max_steps,
                                 PhaseOut* out, unsigned char* lds, size_t lds_bytes, int ib) {
    const int lane = threadIdx.x;
    const int n = (int)g.n;
    const int mask = (1 << ib) - 1;
    const int sh = ib + 1;
    const int stride = (int)g.kstride;
    // keys in banks of at most 32 registers: a uniform dynamic index into one
    // bank is a v_movrel (a larger private array would not stay in registers)
    constexpr int RL = R < 32 ? R : 32;
    int32_t klo[RL], khi[R > 32 ? R - 32 : 1];
#define KEY(r) (*((R <= 32 || (r) < 32) ? &klo[(r) < RL ? (r) : 0] : &khi[(r) >= 32 ? (r) - 32 : 0]))
#define KEY_APPLY(slot, mine, EXPR)                                        \
    if (R <= 32 || (slot) < 32) {                                          \
        const int32_t kv = klo[slot];                                      \
        klo[slot] = kv ^ ((kv ^ (EXPR)) & -(int32_t)(mine));               \
    } else {                                                               \
        const int32_t kv = khi[(slot) - 32];                               \
        khi[(slot) - 32] = kv ^ ((kv ^ (EXPR)) & -(int32_t)(mine));        \
    }
    // padding slots hold the lowest key of node 0: below every real key (the
    // row strengths keep real gains above -2^(30 - ib)) and decoded as node 0
    const int32_t pad = (int32_t)(0x80000000u | ((uint32_t)mask << 1));
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int v = r * 64 + lane;
        if (v < n) {
            const uint8_t xv = xg[v];
            KEY(r) = key_of(xv ? -hg[v] : hg[v], v, xv, ib, mask);
        } else {
            KEY(r) = pad;
        }
    }
    // the neighbour table, zero padded to whole quads (rows zero padded after
    // their entries), is staged in LDS (the host admits only tables that fit;
    // with max_steps = 0 nothing is staged and rows come from global memory)
    __builtin_assume(stride >= 1);
    const int quads = (n * stride + 3) >> 2;
    const bool staged = max_steps > 0 && (size_t)quads * 16 <= lds_bytes;
    if (staged) {
        const int4* src = (const int4*)g.kadj;
        int4* dst = (int4*)lds;
#pragma unroll 12  // 12 loads in flight per lane: a 24 KiB table in two rounds
        for (int q = lane; q < quads; q += 64) dst[q] = src[q];
    }
    __syncthreads();
    // step counters in 32 bits (a descent makes far fewer than 2^31 moves);
    // every pass of the loop scans all n gains, so evals = n (steps + 1)
    const int32_t step_cap = max_steps < INT_MAX ? (int32_t)max_steps : INT_MAX;
    int32_t steps = 0;
    int64_t gsum = 0;
    int32_t bv;
    int bk;
    for (;;) {
        int32_t m0 = KEY(0), m1 = R > 1 ? KEY(1) : INT_MIN, m2 = R > 2 ? KEY(2) : INT_MIN;
#pragma unroll
        for (int r = 3; r + 1 < R; r += 2) {
            if (((r - 3) >> 1) % 3 == 0) m0 = max3i(m0, KEY(r), KEY(r + 1));
            else if (((r - 3) >> 1) % 3 == 1) m1 = max3i(m1, KEY(r), KEY(r + 1));
            else m2 = max3i(m2, KEY(r), KEY(r + 1));
        }
        if constexpr (R > 3 && (R - 3) % 2 == 1) m0 = max(m0, KEY(R - 1));
        const int32_t lm = max3i(m0, m1, m2);
        // neighbour row of this lane's own maximum, loaded while the wave reduces
        // (a lane holding padding only reads row 0 and never wins); all NB
        // loads are issued without branches or waits, entries past the stride
        // repeat entry 0 and are never used.  node = mask - bits = ~bits & mask
        // (the two load paths also keep the loads ahead of the reduction; with
        // max_steps = 0 the rows read are never used)
        int32_t nb[NB];
        const int c = (int)__builtin_amdgcn_ubfe(~(uint32_t)lm, 1u, (uint32_t)ib);
        if (staged) {
            const int32_t* row = (const int32_t*)lds + c * stride;
#pragma unroll
            for (int j = 0; j < NB; ++j) nb[j] = row[j < stride ? j : 0];
        } else {
            const int32_t* row = g.kadj + (int64_t)c * stride;
#pragma unroll
            for (int j = 0; j < NB; ++j) nb[j] = row[j < stride ? j : 0];
        }
        const int32_t m = wave_max_i32(lm);
        bv = m >> sh;
        bk = mask - ((m >> 1) & mask);
        if (!(bv > 0) || steps >= step_cap) break;
        const uint32_t xk = (uint32_t)m & 1u;
        const int sk2 = xk ? -2 : 2;  // 2 s_k before the flip
        const int owner = bk & 63;
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            if (j >= stride) break;
            // a zero padding entry (w = 0) leaves node 0's key unchanged: no test
            const int32_t a = __builtin_amdgcn_readlane(nb[j], owner);
            const int u = a & 0xFFFFFF;
            const int slot = u >> 6;
            const int sk2w = sk2 * (a >> 24);
            // key -= (s_v * 2 s_k w) << sh with s_v = -1 when bit 0 is set:
            // (t ^ m) - m negates t exactly when m = -(bit 0) is -1
            KEY_APPLY(slot, lane == (u & 63),
                      (int32_t)((uint32_t)kv - ((uint32_t)((sk2w ^ __builtin_amdgcn_sbfe(kv, 0u, 1u)) -
                                                            __builtin_amdgcn_sbfe(kv, 0u, 1u)) << sh)));
        }
        // the flipped node: gain negated, spin bit toggled, node bits kept
        const int32_t kk = (int32_t)(((uint32_t)(-bv) << sh) | (((uint32_t)m & ((1u << sh) - 1u)) ^ 1u));
        const int kslot = bk >> 6;
        KEY_APPLY(kslot, lane == owner, kk);
        ++steps;
        gsum += bv;
    }
    if (steps) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int v = r * 64 + lane;
            if (v < n) {
                const int32_t k = KEY(r);
                const int32_t d = k >> sh;
                hg[v] = (k & 1) ? -d : d;
                xg[v] = (uint8_t)(k & 1);
            }
        }
    }
    if (lane == 0) {
        out->gain = (double)bv;
        out->k = bk;
        out->l = -1;
        out->accepts = steps;
        out->evals = (int64_t)n * (steps + 1);
        out->gain_sum = (double)gsum;
    }
}
#undef KEY
#undef KEY_APPLY

template <int R, int NB>
__global__ __launch_bounds__(64) void k_descent_reg(GView<int8_t> g, uint8_t* xg, int32_t* hg, int64_t max_steps,
                                                   PhaseOut* out, size_t lds_bytes, int ib, const int32_t* gate) {
    extern __shared__ __align__(16) unsigned char lds[];
    if (gate && *gate) return;
    descent_reg_body<R, NB>(g, xg, hg, max_steps, out, lds, lds_bytes, ib);
}

// Bit-plane form of the descent for graphs whose nonzero weights are all +-1
// (random regular +-1 graphs, unit weights), n <= 2048 NW, with a neighbour
// table of at most NB entries per row.  Node v's gain is held as
// c_v = (Delta_v + S) >> HS in [0, 2S >> HS] (S: the largest row strength,
// which for +-1 weights is the largest degree; HS = 1 (H >= 1) when every
// degree has the parity of S, so that Delta_v + S is always even; H = 2 when
// also S = 2^NP - 1, so that a flipped node's -Delta_k + S is the complement
// of its planes): bit b of c_v is bit
// (v & 31) of plane word P[b][v >> 11] of lane (v >> 5) & 63, and its spin is
// the same bit of X.  A step's first maximum comes from the planes, most
// significant first: the remaining nodes whose bit b is set replace them
// when there are any (one ballot per plane), and the lowest remaining node is
// the first maximum of pyqrackising/spin_glass_solver.py:45 (a ballot, then
// the lowest bit of one lane's word).  A flip of k changes each neighbour's
// gain by -2 w s_k s_u = +-2: the neighbours to raise and to lower are
// gathered into two bit masks and applied as one bit-sliced increment and
// decrement of (Delta + S) >> 1, and k's own planes are rewritten with
// -Delta_k.  A step costs O(NP + degree) vector instructions for all n gains.
// Rows of at most 3 entries (NB = 3, one word per lane) are held in
// registers: entry j of node v in register v & 31 of bank j of lane v >> 5,
// read by a uniform index and one readlane.  Otherwise, while the step's last
// ballots run, every lane loads the neighbour row of its own lowest remaining
// node from the table in LDS, so the winning lane holds the row of k.  Same
// moves, counters and result as descent_reg_body.
template <int NW, int NP, int NB, int H>
__device__ void descent_plane_body(const GView<int8_t>& g, uint8_t* xg, int32_t* hg, int64_t max_steps,
                                   PhaseOut* out, unsigned char* lds, size_t lds_bytes, int S) {
    constexpr int HS = H ? 1 : 0;  // values halved
    const int lane = threadIdx.x;
    const int n = (int)g.n;
    const int stride = (int)g.kstride;
    uint32_t P[NP][NW], X[NW], V[NW];
    const bool aligned = (((uintptr_t)xg | (uintptr_t)hg) & 15) == 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        const int base = w * 2048 + lane * 32;
        int32_t cv[32];
        uint32_t xb = 0;
        if (aligned && base + 32 <= n) {
            const uint4* xs = (const uint4*)(xg + base);
            const int4* hs = (const int4*)(hg + base);
            const uint4 xa = xs[0], xc = xs[1];
            int4 hq[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) hq[q] = hs[q];
            const uint32_t xw[8] = {xa.x, xa.y, xa.z, xa.w, xc.x, xc.y, xc.z, xc.w};
#pragma unroll
            for (int r = 0; r < 32; ++r) {
                const uint32_t xv = ((xw[r >> 2] >> (8 * (r & 3))) & 0xFFu) != 0;
                const int4 hp = hq[r >> 2];
                const int32_t hv = (r & 3) == 0 ? hp.x : (r & 3) == 1 ? hp.y : (r & 3) == 2 ? hp.z : hp.w;
                cv[r] = ((xv ? -hv : hv) + S) >> HS;
                xb |= xv << r;
            }
        } else {
#pragma unroll
            for (int r = 0; r < 32; ++r) {
                const int v = base + r;
                cv[r] = 0;
                if (v < n) {
                    const uint32_t xv = xg[v] != 0;
                    cv[r] = ((xv ? -hg[v] : hg[v]) + S) >> HS;
                    xb |= xv << r;
                }
            }
        }
        X[w] = xb;
        V[w] = base >= n ? 0u : (base + 32 <= n ? 0xFFFFFFFFu : ((1u << (n - base)) - 1u));
#pragma unroll
        for (int b = 0; b < NP; ++b) {
            uint32_t pb = 0;
#pragma unroll
            for (int r = 0; r < 32; ++r) pb |= (((uint32_t)cv[r] >> b) & 1u) << r;
            P[b][w] = pb;
        }
    }
    // rows of at most 3 entries in registers: three banks of 32, separate
    // arrays so that a uniform dynamic index into each stays a register move.
    // Entries with w = 0 (padding, past the stride, nodes past n) are stored
    // as node 2048, whose lane field (u >> 5 = 64) matches no lane: a no-op
    constexpr bool RR = NB == 3;
    static_assert(!RR || NW == 1, "register rows hold one word of nodes per lane");
    // two planes of (Delta + S) / 2 with S = 3: plane 1 is the positive gains
    constexpr bool SPEC = NW == 1 && NP == 2 && H == 2;
    // register rows are also written to LDS (as the table, v * stride + j),
    // when it fits, for the top-level batch below
    const bool rows_lds = RR && max_steps > 0 && (size_t)n * stride * 4 <= lds_bytes;
    int32_t rr0[RR ? 32 : 1], rr1[RR ? 32 : 1], rr2[RR ? 32 : 1];
    if constexpr (RR) {
        auto rr_set = [&](int j, int r, int32_t v) {
            v = (v >> 24) != 0 ? v : 2048;
            if (j == 0) rr0[r] = v;
            else if (j == 1) rr1[r] = v;
            else rr2[r] = v;
        };
        const int base = lane * 32;
        if (max_steps > 0 && stride == 3 && base + 32 <= n) {
            // 96 consecutive entries in 16-byte pieces (rows start at 12 v bytes)
            const int4* src = (const int4*)(g.kadj + base * 3);
            int4 q[24];
#pragma unroll
            for (int i = 0; i < 24; ++i) q[i] = src[i];
#pragma unroll
            for (int r = 0; r < 32; ++r)
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    const int e = 3 * r + j;
                    const int4 v = q[e >> 2];
                    rr_set(j, r, (e & 3) == 0 ? v.x : (e & 3) == 1 ? v.y : (e & 3) == 2 ? v.z : v.w);
                }
            if (rows_lds) {
#pragma unroll
                for (int i = 0; i < 24; ++i) ((int4*)lds)[lane * 24 + i] = q[i];
            }
        } else {
#pragma unroll
            for (int r = 0; r < 32; ++r)
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    const int v = base + r;
                    const bool in_row = max_steps > 0 && j < stride && v < n;
                    const int32_t a = in_row ? g.kadj[(int64_t)v * stride + j] : 0;
                    rr_set(j, r, a);
                    if (rows_lds && in_row) ((int32_t*)lds)[v * stride + j] = a;
                }
        }
    }
    // otherwise the neighbour table (rows zero padded after their entries, the
    // table to whole quads) is staged in LDS when it fits; with max_steps = 0
    // nothing is staged and no row is used
    __builtin_assume(stride >= 1);
    const int quads = (n * stride + 3) >> 2;
    const bool staged = !RR && max_steps > 0 && (size_t)quads * 16 <= lds_bytes;
    if (staged) {
        const int4* src = (const int4*)g.kadj;
        int4* dst = (int4*)lds;
#pragma unroll 12
        for (int q = lane; q < quads; q += 64) dst[q] = src[q];
    }
    __syncthreads();
    // top-level batch (register rows): bucket words, the (in, out) words of
    // the independent-set rounds, and per-node byte counters
    __shared__ uint32_t top_b[RR ? 64 : 1];
    __shared__ uint64_t top_io[RR ? 64 : 1];
    __shared__ __align__(16) uint32_t top_cnt[RR ? 512 : 4];
    const int32_t step_cap = max_steps < INT_MAX ? (int32_t)max_steps : INT_MAX;
    int32_t steps = 0;
    int64_t gsum = 0;
    int bv = 0, bk = 0;
    if constexpr (RR) {
        // the top-level narrowing of the first step, as in the loop below
        uint32_t M[1] = {V[0]};
        int cm = 0;
#pragma unroll
        for (int b = NP - 1; b >= 0; --b) {
            const uint32_t t = M[0] & P[b][0];
            const bool hit = __ballot(t != 0) != 0;
            M[0] = hit ? t : M[0];
            cm |= hit ? (1 << b) : 0;
        }
        const uint64_t bal = __ballot(M[0] != 0);
        // Top-level batch.  When the largest gain is S, every node B of
        // the bucket has Delta = S >= its row strength, so all its
        // edges are unsatisfied and its flip lowers every neighbour by
        // 2: no gain rises, and the steps take the bucket in index
        // order, each flip dropping its bucket neighbours.  That is
        // the lexicographically first maximal independent set of the
        // bucket's own graph: node v is in it iff none of its lower
        // bucket neighbours is.  It is found in rounds (a node is
        // decided once its lower bucket neighbours are), flipped at
        // once, and each neighbour is lowered by 2 per flipped
        // neighbour.  Same moves and state as the single steps; taken
        // before the first step, when the bucket spans at least 4 lanes
        // and the step cap cannot fall inside it.  Afterwards no gain is
        // S (each bucket node was flipped or lowered), and the single
        // steps go on.
        if (rows_lds && cm == ((2 * S) >> HS) && __popcll(bal) >= 4 && step_cap >= n) {
            // Each lane walks its own bucket nodes (at most the busiest lane's
            // count of steps, a few for a bucket spread over the wave), reading
            // rows from LDS.  Entry j of node v is a lower bucket neighbour
            // when bit r = v & 31 of lbj is set.
            const int32_t* rows = (const int32_t*)lds;
            const uint32_t B = M[0];
            top_b[lane] = B;
            __syncthreads();
            // the (lower bucket neighbour u, bit r of node v) pairs of this
            // lane's bucket nodes, at most 7 per lane, in 8 slots of LDS
            // (top_cnt, not yet in use); a lane with more makes the batch
            // fall back to single steps (nothing has moved yet)
            uint32_t* pairs = top_cnt + lane * 8;
            uint32_t lba = 0u;
            int np = 0;
            for (uint32_t bb = B; __ballot(bb != 0u);) {
                if (bb) {
                    // branch-free: the three entries, then their bucket words,
                    // each group of loads in flight together; a pair is
                    // written at slot np & 7 and kept by advancing np (while
                    // np <= 7 the slot written is never a kept one)
                    const int r = __builtin_ctz(bb);
                    bb &= bb - 1u;
                    const int v = lane * 32 + r;
                    int32_t a[3];
                    uint32_t bw[3];
#pragma unroll
                    for (int j = 0; j < 3; ++j) a[j] = rows[v * stride + min(j, stride - 1)];
#pragma unroll
                    for (int j = 0; j < 3; ++j) bw[j] = top_b[(a[j] & 0x7FF) >> 5];
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        const int u = a[j] & 0x7FF;
                        const bool low = j < stride && (a[j] >> 24) != 0 && u < v && ((bw[j] >> (u & 31)) & 1u);
                        pairs[np & 7] = ((uint32_t)r << 16) | (uint32_t)u;
                        np += low ? 1 : 0;
                        lba |= (uint32_t)low << r;
                    }
                }
            }
            if (__ballot(np > 7) == 0) {
                const uint4 pa = ((const uint4*)pairs)[0], pb = ((const uint4*)pairs)[1];
                const uint32_t pr[8] = {pa.x, pa.y, pa.z, pa.w, pb.x, pb.y, pb.z, pb.w};
                uint32_t in = B & ~lba, out = 0u, und = B & lba;
                while (__ballot(und != 0u)) {
                    top_io[lane] = (uint64_t)in | ((uint64_t)out << 32);
                    __syncthreads();
                    uint32_t hit = 0u, blocked = 0u;
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        // slots past np hold stale values: any node (< 2048) is
                        // a valid index, and the bits are dropped
                        const int u = pr[k] & 0x7FF, r = (pr[k] >> 16) & 31;
                        const uint64_t st = top_io[u >> 5];
                        const uint32_t keep = k < np ? 1u : 0u;
                        hit |= ((((uint32_t)st >> (u & 31)) & keep)) << r;
                        blocked |= ((~(uint32_t)(st >> 32) >> (u & 31)) & keep) << r;
                    }
                    __syncthreads();
                    hit &= und;
                    blocked &= und;
                    out |= hit;
                    const uint32_t join = und & ~hit & ~blocked;
                    in |= join;
                    und &= ~(hit | join);
                }
                // each neighbour's count of flipped neighbours (at most 3)
                uint4* cnt4 = (uint4*)top_cnt;
                cnt4[2 * lane] = make_uint4(0u, 0u, 0u, 0u);
                cnt4[2 * lane + 1] = make_uint4(0u, 0u, 0u, 0u);
                __syncthreads();
                for (uint32_t ii = in; __ballot(ii != 0u);) {
                    if (ii) {
                        const int r = __builtin_ctz(ii);
                        ii &= ii - 1u;
                        const int v = lane * 32 + r;
                        int32_t a[3];
#pragma unroll
                        for (int j = 0; j < 3; ++j) a[j] = rows[v * stride + min(j, stride - 1)];
#pragma unroll
                        for (int j = 0; j < 3; ++j) {
                            // padding entries (and entries past the stride) add 0
                            const int u = a[j] & 0x7FF;
                            const uint32_t add = (j < stride && (a[j] >> 24) != 0) ? 1u << (8 * (u & 3)) : 0u;
                            __hip_atomic_fetch_add(&top_cnt[u >> 2], add, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP);
                        }
                    }
                }
                __syncthreads();
                const uint4 ca = cnt4[2 * lane], cb = cnt4[2 * lane + 1];
                __syncthreads();
                const uint32_t cw[8] = {ca.x, ca.y, ca.z, ca.w, cb.x, cb.y, cb.z, cb.w};
                uint32_t d0 = 0u, d1 = 0u;  // count bits of this lane's 32 nodes
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    // bit 0 of each byte to bits 0..3 (the multiplier's cross
                    // terms stay below bit 24 or leave the word)
                    d0 |= (((cw[q] & 0x01010101u) * 0x01020408u) >> 24) << (4 * q);
                    d1 |= ((((cw[q] >> 1) & 0x01010101u) * 0x01020408u) >> 24) << (4 * q);
                }
                // c -= count (in units of 2 >> HS), a bit-sliced subtraction;
                // the value stays >= 0 (Delta_u >= -deg u), then the flipped
                // nodes take c = 0 (Delta = -S)
                uint32_t bo = 0u;
#pragma unroll
                for (int b = 1 - HS; b < NP; ++b) {
                    const uint32_t sb = b == 1 - HS ? d0 : (b == 2 - HS ? d1 : 0u);
                    const uint32_t pb = P[b][0];
                    P[b][0] = pb ^ sb ^ bo;
                    bo = (~pb & (sb | bo)) | (sb & bo);
                }
#pragma unroll
                for (int b = 0; b < NP; ++b) P[b][0] &= ~in;
                X[0] ^= in;
                const int m = wave_reduce_dpp((int)__popc(in), [](int a, int c) { return a + c; });
                steps = m;
                gsum = (int64_t)S * m;
            }
        }
    }
    for (;;) {
        int L, p, ws = 0, cm = 0;
        int32_t nb[NB];
        {
            uint32_t M[NW];
            uint64_t bal;
            if constexpr (SPEC) {
                // c in 0..3 and Delta > 0 iff c >= 2 (plane 1): both ballots at
                // once.  With no positive gain cm = 0 stops the descent below,
                // and the exact top is found after the loop
                // (two compares, one scalar test feeding both selects, the
                // empty case mapped to lane 63 so the lane stays in range)
                const uint32_t t1 = P[1][0];
                uint32_t mt, t2;
                uint64_t bp, b2, sel;
                __asm__(
                    "v_and_b32 %[t2], %[t1], %[p0]\n\t"
                    "v_cmp_ne_u32_e64 %[b2], 0, %[t2]\n\t"
                    "v_cmp_ne_u32_e64 %[bp], 0, %[t1]\n\t"
                    "s_cmp_lg_u64 %[b2], 0\n\t"
                    "s_cselect_b64 %[bal], %[b2], %[bp]\n\t"
                    "s_cselect_b64 %[sel], -1, 0\n\t"
                    "s_bitset1_b64 %[bal], 63\n\t"
                    "v_cndmask_b32_e64 %[mt], %[t1], %[t2], %[sel]"
                    : [t2] "=&v"(t2), [b2] "=&s"(b2), [bp] "=&s"(bp), [bal] "=&s"(bal), [sel] "=&s"(sel),
                      [mt] "=v"(mt)
                    : [t1] "v"(t1), [p0] "v"(P[0][0])
                    : "scc");
                M[0] = mt;
                cm = b2 ? 3 : (bp ? 2 : 0);
            } else {
#pragma unroll
                for (int w = 0; w < NW; ++w) M[w] = V[w];
#pragma unroll
                for (int b = NP - 1; b >= 0; --b) {
                    uint32_t t[NW], any = 0;
#pragma unroll
                    for (int w = 0; w < NW; ++w) {
                        t[w] = M[w] & P[b][w];
                        any |= t[w];
                    }
                    const bool hit = __ballot(any != 0) != 0;  // wave-uniform
#pragma unroll
                    for (int w = 0; w < NW; ++w) M[w] = hit ? t[w] : M[w];
                    cm |= hit ? (1 << b) : 0;
                }
                // the lowest remaining node: word ws, lane L, bit p.  n >= 1, so
                // M is never empty (a plane only narrows it when its bits meet it)
                bal = __ballot(M[0] != 0);
            }
            uint32_t mw = M[0];
            int cb = 0;
            if constexpr (NW > 1) {
                const uint64_t bal1 = __ballot(M[1] != 0);
                ws = bal ? 0 : 1;
                bal = bal ? bal : bal1;
                cb = M[0] ? 0 : 2048;
                mw = M[0] ? M[0] : M[1];
            }
            // this lane's lowest remaining node and its neighbour row, loaded
            // while the winner is decided (entries past the stride repeat entry
            // 0 and are never used; a lane with nothing left reads a valid row
            // it never uses)
            if constexpr (!RR) {
                const int c = min(cb + lane * 32 + (int)__builtin_ctz(mw | 0x80000000u), n - 1);
                if (staged) {
                    const int32_t* row = (const int32_t*)lds + c * stride;
#pragma unroll
                    for (int j = 0; j < NB; ++j) nb[j] = row[j < stride ? j : 0];
                } else {
                    const int32_t* row = g.kadj + (int64_t)c * stride;
#pragma unroll
                    for (int j = 0; j < NB; ++j) nb[j] = row[j < stride ? j : 0];
                }
            }
            L = (int)__builtin_ctzll(bal);
            p = (int)__builtin_ctz(__builtin_amdgcn_readlane(mw, L) | (SPEC ? 0x80000000u : 0u));
        }
        if constexpr (RR) {
#pragma unroll
            for (int j = 0; j < NB; ++j) nb[j] = j == 0 ? rr0[p] : (j == 1 ? rr1[p] : rr2[p]);  // uniform index
        }
        bk = ws * 2048 + L * 32 + p;
        bv = (cm << HS) - S;
        uint32_t xl = X[0];
        if constexpr (NW > 1) xl = ws ? X[1] : X[0];
        const int32_t xkm = -(int32_t)((__builtin_amdgcn_readlane(xl, L) >> p) & 1u);  // -x_k
        // The flip's new planes are formed before the stop test (bv <= 0 or
        // the step cap, one sign test) is taken, so that test is off the
        // dependency chain; they are kept only when the descent goes on.
        // Neighbours: Delta_u changes by -2 w s_k s_u, i.e. +2 where
        // s_u = -w s_k (u on side 1 when w s_k > 0, side 0 otherwise).
        uint32_t inc[NW], all[NW];
#pragma unroll
        for (int w = 0; w < NW; ++w) inc[w] = all[w] = 0u;
        uint32_t mk1 = 0u;  // register rows: k's own bit, formed with the neighbours' masks
        if constexpr (RR) {
            int32_t a[3];
            int lu[3];
            uint32_t bit[3], m[3], fx[3];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                a[j] = __builtin_amdgcn_readlane(nb[j], L);
                lu[j] = (int)__builtin_amdgcn_ubfe((uint32_t)a[j], 5u, 7u);  // 64: a padding entry
                bit[j] = 1u << (a[j] & 31);
                fx[j] = (uint32_t)((a[j] >> 31) ^ xkm);
            }
            lane_masks4(lane, lu[0], lu[1], lu[2], L, bit[0], bit[1], bit[2], 1u << p, m[0], m[1], m[2], mk1);
            inc[0] = (m[0] & (X[0] ^ fx[0])) | (m[1] & (X[0] ^ fx[1])) | (m[2] & (X[0] ^ fx[2]));
            all[0] = m[0] | m[1] | m[2];
        }
#pragma unroll
        for (int j = 0; j < (RR ? 0 : NB); ++j) {
            if (!RR && j >= stride) break;
            const int32_t a = __builtin_amdgcn_readlane(nb[j], L);
            // lane of u, or 64 for a padding entry (register rows store it as
            // node 2048; table rows are tested here)
            const int lu = RR ? (int)__builtin_amdgcn_ubfe((uint32_t)a, 5u, 7u)
                              : ((a >> 24) != 0 ? ((a >> 5) & 63) : 64);
            const uint32_t m = lane == lu ? 1u << (a & 31) : 0u;
            // fx = 0 when w > 0 and x_k = 0 or w < 0 and x_k = 1, else ~0
            const uint32_t fx = (uint32_t)((a >> 31) ^ xkm);
            if constexpr (NW == 1) {
                inc[0] |= m & (X[0] ^ fx);
                all[0] |= m;
            } else {
                const int u = a & 0xFFFFFF;
#pragma unroll
                for (int w = 0; w < NW; ++w) {
                    const uint32_t mm = (u >> 11) == w ? m : 0u;
                    inc[w] |= mm & (X[w] ^ fx);
                    all[w] |= mm;
                }
            }
        }
        const int cn = (S - bv) >> HS;  // k's new value: (-Delta_k + S) >> HS
        uint32_t Q[NP][NW], Y[NW];
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            const uint32_t mk = RR ? mk1 : ((lane == L && w == ws) ? 1u << p : 0u);
#pragma unroll
            for (int b = 0; b < NP; ++b) {
                if constexpr (H == 2) {
                    Q[b][w] = P[b][w] ^ mk;  // S = 2^NP - 1: S - c is the complement
                } else {
                    const uint32_t nbits = (uint32_t)-((cn >> b) & 1);
                    Q[b][w] = (P[b][w] & ~mk) | (mk & nbits);
                }
            }
            Y[w] = X[w] ^ mk;
            // +2 where inc, -2 where dec: a bit-sliced increment and decrement
            // of (Delta + S) >> 1, planes 1 - HS and up (the value stays in
            // range, so no carry leaves the top plane)
            uint32_t cy = inc[w], bw = all[w] & ~inc[w];
#pragma unroll
            for (int b = 1 - HS; b < NP; ++b) {
                const uint32_t ti = Q[b][w] & cy, td = bw & ~Q[b][w];
                Q[b][w] ^= cy | bw;
                cy = ti;
                bw = td;
            }
        }
        if (((bv - 1) | (step_cap - 1 - steps)) < 0) break;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
#pragma unroll
            for (int b = 0; b < NP; ++b) P[b][w] = Q[b][w];
            X[w] = Y[w];
        }
        ++steps;
        gsum += bv;
    }
    if constexpr (SPEC) {
        if (bv <= 0) {
            // the exact top of the final state: the largest value and its
            // lowest node
            uint32_t M = V[0];
            int cm = 0;
#pragma unroll
            for (int b = NP - 1; b >= 0; --b) {
                const uint32_t t = M & P[b][0];
                const bool hit = __ballot(t != 0) != 0;
                M = hit ? t : M;
                cm |= hit ? (1 << b) : 0;
            }
            const int L = (int)__builtin_ctzll(__ballot(M != 0));
            bk = L * 32 + (int)__builtin_ctz(__builtin_amdgcn_readlane(M, L));
            bv = (cm << HS) - S;
        }
    }
    if (steps) {
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            const int base = w * 2048 + lane * 32;
            int32_t hv[32];
            uint32_t xw[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int r = 0; r < 32; ++r) {
                int32_t cr = 0;
#pragma unroll
                for (int b = 0; b < NP; ++b) cr |= (int32_t)((P[b][w] >> r) & 1u) << b;
                const uint32_t xv = (X[w] >> r) & 1u;
                const int32_t d = (cr << HS) - S;
                hv[r] = xv ? -d : d;
                xw[r >> 2] |= xv << (8 * (r & 3));
            }
            if (aligned && base + 32 <= n) {
                int4* hs = (int4*)(hg + base);
#pragma unroll
                for (int q = 0; q < 8; ++q) hs[q] = make_int4(hv[4 * q], hv[4 * q + 1], hv[4 * q + 2], hv[4 * q + 3]);
                uint4* xs = (uint4*)(xg + base);
                xs[0] = make_uint4(xw[0], xw[1], xw[2], xw[3]);
                xs[1] = make_uint4(xw[4], xw[5], xw[6], xw[7]);
            } else {
#pragma unroll
                for (int r = 0; r < 32; ++r) {
                    const int v = base + r;
                    if (v < n) {
                        hg[v] = hv[r];
                        xg[v] = (uint8_t)((xw[r >> 2] >> (8 * (r & 3))) & 1u);
                    }
                }
            }
        }
    }
    if (lane == 0) {
        out->gain = (double)bv;
        out->k = bk;
        out->l = -1;
        out->accepts = steps;
        out->evals = (int64_t)n * (steps + 1);
        out->gain_sum = (double)gsum;
    }
}

template <int NW, int NP, int NB, int H>
__global__ __launch_bounds__(64) void k_descent_plane(GView<int8_t> g, uint8_t* xg, int32_t* hg, int64_t max_steps,
                                                     PhaseOut* out, size_t lds_bytes, int S, const int32_t* gate) {
    extern __shared__ __align__(16) unsigned char lds[];
    if (gate && *gate) return;
    descent_plane_body<NW, NP, NB, H>(g, xg, hg, max_steps, out, lds, lds_bytes, S);
}

// Registers per lane of the register-resident descent for n nodes with a
// neighbour table of `stride` entries per row, or 0 when it does not apply.
static int reg_descent_regs(int64_t n, int64_t stride) {
    if (stride <= 0 || stride > REG_NB) return 0;
    for (int r = 8; r <= 64; r *= 2)
        if (n <= 64 * r) return r;
    return 0;
}

static size_t reg_table_bytes(const Graph& g) {
    return (size_t)((g.n * g.key_adj_stride + 3) >> 2) * 16;
}

// variant 9: the register-resident descent without the bit-plane form
static int reg_descent_enabled() { return descent_variant() == 0 || descent_variant() == 9; }

template <int R, int NB>
__global__ __launch_bounds__(64) void k_descent_reg(GView<int8_t> g, uint8_t* xg, int32_t* hg, int64_t max_steps,
                                                   PhaseOut* out, size_t lds_bytes, int ib, const int32_t* gate);

// LDS available to the register-resident kernels (their neighbour table must fit)
static size_t reg_lds_cap() {
    static const size_t cap = lds_limit((const void*)k_descent_reg<64, 16>);
    return cap;
}

int reg_descent_plan(const Graph& g, size_t* table_bytes) {
    if (!reg_descent_enabled() || key_shift(g) <= 0 || !g.key_adj) return 0;
    const int r = reg_descent_regs(g.n, g.key_adj_stride);
    if (!r || reg_table_bytes(g) > reg_lds_cap()) return 0;
    if (table_bytes) *table_bytes = reg_table_bytes(g);
    return r;
}

// neighbour-row width of the register-resident descent: the stride rounded up to 4, 8 or 16
static int reg_nb(int64_t stride) { return stride <= 4 ? 4 : (stride <= 8 ? 8 : 16); }

// Bit-plane descent classes: S <= 3, 7, 15, 16 (a +-1 graph's degree is its
// strength, so rows of at most 3 (4 with two words per lane), 8, 16, 16
// entries), each with 3 - H, 4 - H, 5 - H, 6 - H planes for values up to
// 2S >> H.
static int plane_class(int64_t S) { return S <= 3 ? 0 : (S <= 7 ? 1 : (S <= 15 ? 2 : 3)); }

template <int NW, int NP, int NB, int H>
__global__ __launch_bounds__(64) void k_descent_plane(GView<int8_t> g, uint8_t* xg, int32_t* hg, int64_t max_steps,
                                                     PhaseOut* out, size_t lds_bytes, int S, const int32_t* gate);

static size_t plane_lds_cap() {
    static const size_t cap = lds_limit((const void*)k_descent_plane<2, 6, 16, 0>);
    return cap;
}

int plane_descent_plan(const Graph& g, size_t* table_bytes) {
    if (descent_variant() != 0 || g.kind != PQI_WEIGHTS_I8 || !g.unit_abs || !g.key_adj || g.n < 1 || g.n > 4096 ||
        g.max_strength < 0 || g.max_strength > 16 || g.key_adj_stride > g.max_strength)
        return 0;
    if (reg_table_bytes(g) > plane_lds_cap()) return 0;
    if (table_bytes) *table_bytes = reg_table_bytes(g);
    return g.n <= 2048 ? 1 : 2;
}

// whether every degree of g has the parity of S (then Delta + S is even)
int plane_half(const Graph& g, int64_t S) { return g.degree_parity >= 0 && g.degree_parity == (int)(S & 1); }

// H of the bit-plane kernels: 2 when the values are halved and S = 2^NP - 1
// (S = 3, 7, 15 with NP = 2, 3, 4), 1 when only halved, 0 otherwise
static int plane_h(int half, int64_t S) {
    if (!half) return 0;
    const int np = plane_class(S) + 2;
    return S == (int64_t(1) << np) - 1 ? 2 : 1;
}

#define PQI_PLANE_H(NW, CLS, NB, H, KERNEL, ...)                         \
    if ((H) == 2) KERNEL<NW, CLS + 2, NB, 2>(__VA_ARGS__);              \
    else if ((H) == 1) KERNEL<NW, CLS + 2, NB, 1>(__VA_ARGS__);         \
    else KERNEL<NW, CLS + 3, NB, 0>(__VA_ARGS__);
#define PQI_PLANE_CLS(NW, CLS, H, KERNEL, ...)                           \
    switch (CLS) {                                                      \
        case 0: PQI_PLANE_H(NW, 0, (NW == 1 ? 3 : 4), H, KERNEL, __VA_ARGS__) break; \
        case 1: PQI_PLANE_H(NW, 1, 8, H, KERNEL, __VA_ARGS__) break;    \
        case 2: PQI_PLANE_H(NW, 2, 16, H, KERNEL, __VA_ARGS__) break;   \
        default: PQI_PLANE_H(NW, 3, 16, H, KERNEL, __VA_ARGS__) break;  \
    }
#define PQI_PLANE_DISPATCH(NW, CLS, H, KERNEL, ...)                      \
    switch (NW) {                                                       \
        case 1: PQI_PLANE_CLS(1, CLS, H, KERNEL, __VA_ARGS__); break;   \
        default: PQI_PLANE_CLS(2, CLS, H, KERNEL, __VA_ARGS__); break;  \
    }

template <int NW, int NP, int NB, int H>
static void launch_plane(const GView<int8_t>& v, uint8_t* x, int32_t* h, int64_t max_steps, PhaseOut* out,
                         size_t table, int S, hipStream_t s, hipEvent_t e0, hipEvent_t e1, const int32_t* gate) {
    hipExtLaunchKernelGGL((k_descent_plane<NW, NP, NB, H>), dim3(1), dim3(64), (uint32_t)table, s, e0, e1, 0, v,
                          x, h, max_steps, out, table, S, gate);
}

#define PQI_REG_NB(R, NB, KERNEL, ...)                                   \
    switch (NB) {                                                       \
        case 4: KERNEL<R, 4>(__VA_ARGS__); break;                       \
        case 8: KERNEL<R, 8>(__VA_ARGS__); break;                       \
        default: KERNEL<R, 16>(__VA_ARGS__); break;                     \
    }
#define PQI_REG_DISPATCH(R, NB, KERNEL, ...)                             \
    switch (R) {                                                        \
        case 8: PQI_REG_NB(8, NB, KERNEL, __VA_ARGS__); break;          \
        case 16: PQI_REG_NB(16, NB, KERNEL, __VA_ARGS__); break;        \
        case 32: PQI_REG_NB(32, NB, KERNEL, __VA_ARGS__); break;        \
        default: PQI_REG_NB(64, NB, KERNEL, __VA_ARGS__); break;        \
    }

template <int R, int NB>
static void launch_reg(const GView<int8_t>& v, uint8_t* x, int32_t* h, int64_t max_steps, PhaseOut* out,
                       size_t table, int ib, hipStream_t s, hipEvent_t e0, hipEvent_t e1, const int32_t* gate) {
    static const size_t limit = lds_limit((const void*)k_descent_reg<R, NB>);
    (void)limit;  // the plan admits only tables within it
    hipExtLaunchKernelGGL((k_descent_reg<R, NB>), dim3(1), dim3(64), (uint32_t)table, s, e0, e1, 0, v, x, h,
                          max_steps, out, table, ib, gate);
}

// Key width for the packed-key descent of g, or 0 when it does not apply.
int key_shift(const Graph& g) {
    if (g.kind != PQI_WEIGHTS_I8 || g.max_strength < 0 || g.n > (1 << 24)) return 0;
    int ib = 1;
    while ((int64_t(1) << ib) < g.n) ++ib;
    if (ib > 29 || g.max_strength >= (int64_t(1) << (30 - ib))) return 0;
    return ib;
}

// Dynamic LDS for the packed-key descent of g: the first staged layout that
// fits in `limit`, else the keys alone.
size_t key_lds_bytes(const Graph& g, size_t limit) {
    const bool lists = g.use_adj || g.W == nullptr;
    const KeyLds L = key_lds_layout(g.n, lists ? g.nnz : 0, g.key_adj_stride, lists);
    switch (key_stage_mode(L, limit)) {
        case 1: return L.fixed;
        case 2: return L.csr;
        default: return L.base;
    }
}

// Sparse-graph variant with a two-level argmax: LDS keeps, for every group of
// 64 consecutive nodes, the group's first maximum gain and its node.  A step
// takes the first maximum over the group maxima (ties -> lowest node, since
// groups are in node order), and a flip of k only recomputes the groups of k
// and of its neighbours.  Same selection rule and result as the full scan.
template <typename FT>
__device__ __forceinline__ void group_max(const FT* D, int64_t n, int64_t grp, FT* gmax, int32_t* gidx) {
    const int lane = threadIdx.x;
    const int64_t v = grp * 64 + lane;
    FT bv = v < n ? D[v] : lowest<FT>();
    int bk = v < n ? (int)v : INT_MAX;
    wave_argmax_dpp(bv, bk);
    if (lane == 0) {
        gmax[grp] = bv;
        gidx[grp] = bk;
    }
}

template <typename WT, typename FT>
__device__ void descent_wave2_body(const GView<WT>& g, uint8_t* xg, FT* hg, int64_t max_steps, PhaseOut* out,
                                   unsigned char* lds) {
    const int lane = threadIdx.x;
    const int n = (int)g.n;
    const int ng = (n + 63) >> 6;
    size_t off = 0;
    FT* D = (FT*)(lds + off);
    off += ((size_t)n * sizeof(FT) + 15) & ~(size_t)15;
    FT* gmax = (FT*)(lds + off);
    off += ((size_t)ng * sizeof(FT) + 15) & ~(size_t)15;
    int32_t* gidx = (int32_t*)(lds + off);
    off += ((size_t)ng * 4 + 15) & ~(size_t)15;
    uint8_t* x = lds + off;
    for (int v = lane; v < n; v += 64) {
        const uint8_t xv = xg[v];
        x[v] = xv;
        D[v] = xv ? -hg[v] : hg[v];
    }
    __syncthreads();
    for (int grp = 0; grp < ng; ++grp) group_max(D, n, grp, gmax, gidx);
    __syncthreads();
    int64_t steps = 0, evals = 0;
    double gsum = 0.0;
    FT bv;
    int bk;
    for (;;) {
        bv = lowest<FT>();
        bk = INT_MAX;
        for (int t = lane; t < ng; t += 64) {
            const FT d = gmax[t];
            const int k = gidx[t];
            if (d > bv || (d == bv && k < bk)) { bv = d; bk = k; }
        }
        wave_argmax_dpp(bv, bk);
        evals += n;
        if (!(bv > 0) || steps >= max_steps) break;
        const int sk = x[bk] ? -1 : 1;
        __syncthreads();
        const int32_t e0 = g.ap[bk], e1 = g.ap[bk + 1];
        // first 64 neighbours stay in registers for the group refresh below
        int nb0 = -1;
        for (int32_t c = e0; c < e1; c += 64) {
            const int32_t e = c + lane;
            if (e < e1) {
                const int v = g.ai[e];
                if (c == e0) nb0 = v;
                const int sv = x[v] ? -1 : 1;
                D[v] -= (FT)(2 * sk * sv) * (FT)g.aw[e];
            }
        }
        __syncthreads();
        if (lane == 0) {
            D[bk] = -D[bk];
            x[bk] ^= 1u;
        }
        __syncthreads();
        const int gk = bk >> 6;
        group_max(D, n, gk, gmax, gidx);
        for (int32_t c = e0; c < e1; c += 64) {
            const int nbv = c == e0 ? nb0 : (c + lane < e1 ? g.ai[c + lane] : -1);
            const int cnt = e1 - c < 64 ? e1 - c : 64;
            int prev = gk;
            for (int t = 0; t < cnt; ++t) {
                const int grp = __shfl(nbv, t, 64) >> 6;
                if (grp != prev && grp != gk) group_max(D, n, grp, gmax, gidx);
                prev = grp;
            }
        }
        __syncthreads();
        ++steps;
        gsum += (double)bv;
    }
    if (steps) {
        for (int v = lane; v < n; v += 64) {
            const uint8_t xv = x[v];
            hg[v] = xv ? -D[v] : D[v];
            xg[v] = xv;
        }
    }
    if (lane == 0) {
        out->gain = (double)bv;
        out->k = bk;
        out->l = -1;
        out->accepts = steps;
        out->evals = evals;
        out->gain_sum = gsum;
    }
}

template <typename WT, typename FT>
__global__ __launch_bounds__(64) void k_descent_wave2(GView<WT> g, uint8_t* xg, FT* hg, int64_t max_steps,
                                                     PhaseOut* out, const int32_t* gate) {
    extern __shared__ __align__(16) unsigned char lds[];
    if (gate && *gate) return;
    descent_wave2_body<WT, FT>(g, xg, hg, max_steps, out, lds);
}

template <typename FT>
static size_t descent2_lds(int64_t n) {
    const int64_t ng = (n + 63) >> 6;
    return (((size_t)n * sizeof(FT) + 15) & ~(size_t)15) + (((size_t)ng * sizeof(FT) + 15) & ~(size_t)15) +
           (((size_t)ng * 4 + 15) & ~(size_t)15) + (size_t)n;
}

// mode 0: descent_wave_body; 1: the same with the row pointers staged in
// LDS; 2: descent_round_body (sparse graphs)
template <typename WT, typename FT>
__global__ __launch_bounds__(64) void k_descent_wave(GView<WT> g, uint8_t* xg, FT* hg, int64_t max_steps,
                                                    PhaseOut* out, const int32_t* gate, int mode) {
    extern __shared__ __align__(16) unsigned char lds[];
    if (gate && *gate) return;
    descent_wave_body<WT, FT>(g, xg, hg, max_steps, out, lds, mode != 0 && g.use_adj);
}

// the round descent in a kernel of its own (its register allocation is not
// shared with the full-scan body)
template <typename WT, typename FT>
__global__ __launch_bounds__(64) void k_descent_round(GView<WT> g, uint8_t* xg, FT* hg, int64_t max_steps,
                                                     PhaseOut* out, const int32_t* gate) {
    extern __shared__ __align__(16) unsigned char lds[];
    if (gate && *gate) return;
    descent_round_body<WT, FT>(g, xg, hg, max_steps, out, lds);
}

// the round descent applies to sparse graphs whose gains, spins, row pointers
// and candidate rows fit in LDS (node indices below 2^22: they share an int
// with the candidate slot)
template <typename WT, typename FT>
static bool round_fits(int64_t n, size_t limit) {
    return descent_variant() != 8 && n < (int64_t(1) << 22) && round_lds_layout<WT, FT>(n).total <= limit;
}

// LDS of descent_wave_body: gains and spins, plus the row pointers when staged
template <typename FT>
static size_t wave_lds(int64_t n, bool stage_ap) {
    return (((size_t)n * sizeof(FT) + 15) & ~(size_t)15) + (((size_t)n + 15) & ~(size_t)15) +
           (stage_ap ? (size_t)(n + 1) * 4 : 0);
}

// The descent launches go through hipExtLaunchKernel so that optional start
// and stop events time the kernel itself (no events: a plain launch).
template <typename WT>
static void descent_impl(const Graph& g, uint8_t* x, void* h, int64_t max_steps, PhaseOut* out, hipStream_t s,
                         hipEvent_t e0, hipEvent_t e1, const int32_t* gate) {
    using FT = typename FieldOf<WT>::T;
    if constexpr (sizeof(WT) == 1) {
        size_t table = 0;
        const int nw = plane_descent_plan(g, &table);
        if (nw) {
            const int S = (int)g.max_strength;
            PQI_PLANE_DISPATCH(nw, plane_class(S), plane_h(plane_half(g, S), S), launch_plane, view<int8_t>(g), x,
                               (int32_t*)h, max_steps, out, table, S, s, e0, e1, gate);
            return;
        }
        const int regs = reg_descent_plan(g, &table);
        if (regs) {
            PQI_REG_DISPATCH(regs, reg_nb(g.key_adj_stride), launch_reg, view<int8_t>(g), x, (int32_t*)h, max_steps,
                             out, table, key_shift(g), s, e0, e1, gate);
            return;
        }
        static const size_t klimit = lds_limit((const void*)k_descent_key);
        const int ib = key_shift(g);
        const size_t kl = key_lds_bytes(g, klimit);
        if (ib > 0 && kl <= klimit && descent_variant() != 3) {
            hipExtLaunchKernelGGL(k_descent_key, dim3(1), dim3(64), (uint32_t)kl, s, e0, e1, 0, view<int8_t>(g), x,
                                  (int32_t*)h, max_steps, out, kl, ib, gate);
            return;
        }
    }
    static const size_t limit = std::min(lds_limit((const void*)k_descent_wave<WT, FT>),
                                         lds_limit((const void*)k_descent_round<WT, FT>));
    static const size_t limit2 = lds_limit((const void*)k_descent_wave2<WT, FT>);
    const bool sparse = (g.use_adj || g.W == nullptr) && descent_variant() == 2;
    const bool stage_ap = g.use_adj && descent_variant() != 7 && wave_lds<FT>(g.n, true) <= limit;
    const bool rounds = g.use_adj && round_fits<WT, FT>(g.n, limit);
    const size_t lds = rounds ? round_lds_layout<WT, FT>(g.n).total : wave_lds<FT>(g.n, stage_ap);
    if (sparse && descent2_lds<FT>(g.n) <= limit2)
        hipExtLaunchKernelGGL(k_descent_wave2<WT, FT>, dim3(1), dim3(64), (uint32_t)descent2_lds<FT>(g.n), s, e0, e1,
                              0, view<WT>(g), x, (FT*)h, max_steps, out, gate);
    else if (rounds)
        hipExtLaunchKernelGGL(k_descent_round<WT, FT>, dim3(1), dim3(64), (uint32_t)lds, s, e0, e1, 0, view<WT>(g), x,
                              (FT*)h, max_steps, out, gate);
    else if (lds <= limit)
        hipExtLaunchKernelGGL(k_descent_wave<WT, FT>, dim3(1), dim3(64), (uint32_t)lds, s, e0, e1, 0, view<WT>(g), x,
                              (FT*)h, max_steps, out, gate, stage_ap ? 1 : 0);
    else
        hipExtLaunchKernelGGL(k_descent<WT, FT>, dim3(1), dim3(1024), 0, s, e0, e1, 0, view<WT>(g), x, (FT*)h,
                              max_steps, out, gate);
}

void launch_descent(const Graph& g, uint8_t* x, void* h, int64_t max_steps, PhaseOut* out, hipStream_t s,
                    hipEvent_t e0, hipEvent_t e1, const int32_t* gate) {
    PQI_DISPATCH(g, descent_impl, g, x, h, max_steps, out, s, e0, e1, gate);
}

// Best single flip only: one scan of Delta_k = s_k h_k over the state in
// global memory by a 1024-thread work group (no staging, no moves).
template <typename WT, typename FT>
__global__ __launch_bounds__(1024) void k_single_best(GView<WT> g, uint8_t* x, FT* h, PhaseOut* out) {
    descent_body<WT, FT>(g, x, h, 0, out);
}

template <typename WT>
static void single_best_impl(const Graph& g, const uint8_t* x, const void* h, PhaseOut* out, hipStream_t s) {
    using FT = typename FieldOf<WT>::T;
    hipLaunchKernelGGL((k_single_best<WT, FT>), dim3(1), dim3(1024), 0, s, view<WT>(g), (uint8_t*)x, (FT*)h, out);
}

void launch_single_best(const Graph& g, const uint8_t* x, const void* h, PhaseOut* out, hipStream_t s) {
    PQI_DISPATCH(g, single_best_impl, g, x, h, out, s);
}

// ---------------------------------------------------------------------------
// Flip application (host-selected moves)
// ---------------------------------------------------------------------------

template <typename WT, typename FT>
__device__ void flip_body(const GView<WT>& g, uint8_t* x, FT* h, int64_t k, int64_t l) {
    if (k < 0) return;
    int s_old = x[k] ? -1 : 1;
    __syncthreads();
    flip_update(g, x, h, k, s_old, threadIdx.x, blockDim.x);
    __syncthreads();
    if (l >= 0) {
        s_old = x[l] ? -1 : 1;
        __syncthreads();
        flip_update(g, x, h, l, s_old, threadIdx.x, blockDim.x);
    }
}

template <typename WT, typename FT>
__global__ __launch_bounds__(1024) void k_flip(GView<WT> g, uint8_t* x, FT* h, int64_t k, int64_t l) {
    flip_body<WT, FT>(g, x, h, k, l);
}

template <typename WT>
static void flip_impl(const Graph& g, uint8_t* x, void* h, int64_t k, int64_t l, hipStream_t s) {
    using FT = typename FieldOf<WT>::T;
    hipLaunchKernelGGL((k_flip<WT, FT>), dim3(1), dim3(1024), 0, s, view<WT>(g), x, (FT*)h, k, l);
}

void launch_flip(const Graph& g, uint8_t* x, void* h, int64_t k, int64_t l, hipStream_t s) {
    PQI_DISPATCH(g, flip_impl, g, x, h, k, l, s);
}

// ---------------------------------------------------------------------------
// Double flips
// ---------------------------------------------------------------------------

// Row-major pair index s -> (i, j), i < j (pyqrackising/spin_glass_solver.py:66-76
// walks the rows one by one).  Row i starts at base(i) = i(2n-i-1)/2; i is the
// root of base(i) <= s, taken in FP64 and corrected by exact integer checks.
__device__ __forceinline__ int64_t pair_base(int64_t i, int64_t n) { return i * (2 * n - i - 1) / 2; }

__device__ __forceinline__ void unrank_pair(int64_t s, int64_t n, int64_t& i, int64_t& j) {
    const double b = 2.0 * (double)n - 1.0;
    int64_t r = (int64_t)((b - sqrt(b * b - 8.0 * (double)s)) * 0.5);
    if (r < 0) r = 0;
    if (r > n - 2) r = n - 2;
    while (r > 0 && pair_base(r, n) > s) --r;
    while (r < n - 2 && pair_base(r + 1, n) <= s) ++r;
    i = r;
    j = s - pair_base(r, n) + r + 1;
}

template <typename WT, typename FT>
__device__ __forceinline__ FT pair_gain(const GView<WT>& g, const uint8_t* x, const FT* h, int64_t i, int64_t j) {
    const int si = x[i] ? -1 : 1, sj = x[j] ? -1 : 1;
    return (FT)si * h[i] + (FT)sj * h[j] - (FT)(2 * si * sj) * (FT)weight_of(g, i, j);
}

// Reference CPU semantics (D2): thread t of thread_count evaluates the last
// pair index t + tb*floor((C-1-t)/tb) of its strided loop; argmax over t.
template <typename WT, typename FT>
__device__ void double_ref_body(const GView<WT>& g, const uint8_t* x, const FT* h, int64_t thread_count,
                                PhaseOut* out) {
    __shared__ FT sv[16];
    __shared__ int64_t sk[16];
    const int64_t n = g.n, C = n * (n - 1) / 2;
    const int64_t tb = (C + thread_count - 1) / thread_count;
    FT bv = lowest<FT>();
    int64_t bt = LLONG_MAX;
    for (int64_t t = threadIdx.x; t < thread_count; t += blockDim.x) {
        if (t >= C) continue;
        int64_t s = t + tb * ((C - 1 - t) / tb), i, j;
        unrank_pair(s, n, i, j);
        FT d = pair_gain(g, x, h, i, j);
        if (better(d, t, bv, bt)) { bv = d; bt = t; }
    }
    block_argmax(bv, bt, sv, sk);
    if (threadIdx.x == 0) {
        int64_t s = bt + tb * ((C - 1 - bt) / tb), i, j;
        unrank_pair(s, n, i, j);
        out->gain = (double)bv;
        out->k = i;
        out->l = j;
        out->accepts = 0;
        out->evals = thread_count < C ? thread_count : C;
        out->gain_sum = 0.0;
    }
}

template <typename WT, typename FT>
__global__ __launch_bounds__(1024) void k_double_ref(GView<WT> g, const uint8_t* x, const FT* h,
                                                    int64_t thread_count, PhaseOut* out) {
    double_ref_body<WT, FT>(g, x, h, thread_count, out);
}

// Exhaustive: first maximum over all pairs in pair-index order (kernels.cl:186-231).
// Stage 1: each work group scans rows i = blockIdx.x, +gridDim.x, ...
template <typename WT, typename FT>
__global__ __launch_bounds__(256) void k_double_exh(GView<WT> g, const uint8_t* x, const FT* h, FT* pv,
                                                   int64_t* ps) {
    __shared__ FT sv[4];
    __shared__ int64_t sk[4];
    const int64_t n = g.n;
    FT bv = lowest<FT>();
    int64_t bs = LLONG_MAX;
    for (int64_t i = blockIdx.x; i < n - 1; i += gridDim.x) {
        const int64_t base = i * (2 * n - i - 1) / 2 - i - 1;  // pair index of (i, j) = base + j
        const int si = x[i] ? -1 : 1;
        const FT di = (FT)si * h[i];
        for (int64_t j = i + 1 + threadIdx.x; j < n; j += blockDim.x) {
            const int sj = x[j] ? -1 : 1;
            FT d = di + (FT)sj * h[j];
            WT w = g.W ? g.W[i * g.ld + j] : weight_of(g, i, j);
            d -= (FT)(2 * si * sj) * (FT)w;
            int64_t s = base + j;
            if (better(d, s, bv, bs)) { bv = d; bs = s; }
        }
    }
    block_argmax(bv, bs, sv, sk);
    if (threadIdx.x == 0) {
        pv[blockIdx.x] = bv;
        ps[blockIdx.x] = bs;
    }
}

template <typename FT>
__global__ __launch_bounds__(1024) void k_double_exh_final(const FT* pv, const int64_t* ps, int nparts, int64_t n,
                                                          PhaseOut* out) {
    __shared__ FT sv[16];
    __shared__ int64_t sk[16];
    FT bv = lowest<FT>();
    int64_t bs = LLONG_MAX;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x)
        if (better(pv[i], ps[i], bv, bs)) { bv = pv[i]; bs = ps[i]; }
    block_argmax(bv, bs, sv, sk);
    if (threadIdx.x == 0) {
        int64_t i = -1, j = -1;
        if (bs != LLONG_MAX) unrank_pair(bs, n, i, j);
        out->gain = (double)bv;
        out->k = i;
        out->l = j;
        out->accepts = 0;
        out->evals = n * (n - 1) / 2;
        out->gain_sum = 0.0;
    }
}

static const int kExhParts = 1024;

size_t double_scratch_bytes(const Graph&) { return (size_t)kExhParts * 16; }

template <typename WT>
static void double_impl(const Graph& g, const uint8_t* x, const void* h, int64_t thread_count, int mode,
                        PhaseOut* out, void* scratch, hipStream_t s) {
    using FT = typename FieldOf<WT>::T;
    GView<WT> v = view<WT>(g);
    if (mode == PQI_DOUBLE_REFERENCE) {
        hipLaunchKernelGGL((k_double_ref<WT, FT>), dim3(1), dim3(1024), 0, s, v, x, (const FT*)h, thread_count, out);
    } else {
        int parts = (int)std::min<int64_t>(kExhParts, std::max<int64_t>(1, g.n - 1));
        FT* pv = (FT*)scratch;
        int64_t* ps = (int64_t*)((char*)scratch + (size_t)kExhParts * 8);
        hipLaunchKernelGGL((k_double_exh<WT, FT>), dim3(parts), dim3(256), 0, s, v, x, (const FT*)h, pv, ps);
        hipLaunchKernelGGL((k_double_exh_final<FT>), dim3(1), dim3(1024), 0, s, pv, ps, parts, g.n, out);
    }
}

void launch_double_best(const Graph& g, const uint8_t* x, const void* h, int64_t thread_count, int mode,
                        PhaseOut* out, void* scratch, hipStream_t s) {
    PQI_DISPATCH(g, double_impl, g, x, h, thread_count, mode, out, scratch, s);
}

// ---------------------------------------------------------------------------
// Gray seeds (pyqrackising/spin_glass_solver.py:95-123 with the mutation of
// maxcut_tfim_util.py:740-749): seed b mutates theta at positions
// n - offset - 1 - j for the set bits j of gray(b mod block_size), offset =
// 64*floor(b / block_size); a negative position wraps once by +n.
// ---------------------------------------------------------------------------

__device__ __forceinline__ int64_t seed_pos(int64_t n, int64_t offset, int j) {
    int64_t idx = n - offset - 1 - j;
    while (idx < 0) idx += n;
    return idx;
}

// Word w of seed row b, t = b * wpr + w (one thread per word).
__device__ __forceinline__ void seed_bits_body(int64_t t, const uint8_t* x, int64_t n, int64_t wpr, int64_t bs,
                                               int64_t count, uint32_t* bits, unsigned long long* qzero,
                                               int32_t* rzero) {
    if (t >= count * wpr) return;
    const int64_t b = t / wpr, w = t % wpr;
    if (w == 0) {
        // this seed's MFMA row total and rank count start at zero
        if (qzero) qzero[b] = 0ull;
        if (rzero) rzero[b] = 0;
    }
    // the state's word w: 4 bytes of 0/1 -> 4 bits by one multiply, when the
    // 32 bytes lie inside x and x is 16-byte aligned
    uint32_t word = 0;
    if (w * 32 + 32 <= n && ((uintptr_t)x & 15) == 0) {
        const uint4* x4 = (const uint4*)(x + w * 32);
        const uint4 lo = x4[0], hi = x4[1];
        const uint32_t u[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
#pragma unroll
        for (int q = 0; q < 8; ++q) word |= ((u[q] * 0x01020408u) >> 24 & 0xFu) << (4 * q);
    } else {
        for (int q = 0; q < 32; ++q) {
            const int64_t v = w * 32 + q;
            if (v < n && x[v]) word |= 1u << q;
        }
    }
    const uint64_t i = (uint64_t)(b % bs);
    const uint64_t gray = i ^ (i >> 1);
    const int64_t offset = (b / bs) << 6;
    const int64_t base = n - offset - 1;  // position of gray bit 0
    if (n >= 64 && base >= 63) {
        // no wrap: bit j flips position base - j; with r = bit-reverse(gray),
        // position p takes bit p + 63 - base of r
        const uint64_t r = __brevll(gray);
        const int64_t sh = w * 32 + 63 - base;
        uint32_t mask = 0;
        if (sh >= 0 && sh < 64) mask = (uint32_t)(r >> sh);
        else if (sh < 0 && sh > -32) mask = (uint32_t)(r << (-sh));
        word ^= mask;
    } else {
        for (int j = 0; j < 64 && j < n; ++j) {
            if (!((gray >> j) & 1u)) continue;
            const int64_t idx = seed_pos(n, offset, j);
            if ((idx >> 5) == w) word ^= 1u << (idx & 31);
        }
    }
    bits[b * wpr + w] = word;
}

__global__ __launch_bounds__(256) void k_seed_bits(const uint8_t* x, int64_t n, int64_t wpr, int64_t bs,
                                                   int64_t count, uint32_t* bits, unsigned long long* qzero,
                                                   int32_t* rzero, const int32_t* gate) {
    if (gate && *gate) return;
    seed_bits_body((int64_t)blockIdx.x * blockDim.x + threadIdx.x, x, n, wpr, bs, count, bits, qzero, rzero);
}

void launch_seed_bits(const uint8_t* x, int64_t n, int64_t wpr, int64_t bs, int64_t count, uint32_t* bits,
                      hipStream_t s, unsigned long long* qzero, const int32_t* gate, int32_t* rzero) {
    int64_t threads = count * wpr;
    hipLaunchKernelGGL(k_seed_bits, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, x, n, wpr, bs, count,
                       bits, qzero, rzero, gate);
}

// Stable top k of the int8 MFMA seeds straight from their row totals: a larger
// cut difference (q_base - q_b) / 4 is a smaller q_b, so seed a's rank counts
// the seeds with a smaller q, or an equal q and a lower index -- the order of
// k_rank over the cut differences.  Comparisons are split over gridDim.y work
// groups (integer atomics on counts zeroed by k_seed_bits).
__global__ __launch_bounds__(256) void k_rank_q(const unsigned long long* qi, int64_t count, int64_t span,
                                               int32_t* rank) {
    __shared__ long long tile[256];
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const long long qa = a < count ? (long long)qi[a] : 0;
    const int64_t lo = (int64_t)blockIdx.y * span;
    const int64_t hi = lo + span < count ? lo + span : count;
    int32_t r = 0;
    for (int64_t t0 = lo; t0 < hi; t0 += 256) {
        __syncthreads();
        if (t0 + threadIdx.x < hi) tile[threadIdx.x] = (long long)qi[t0 + threadIdx.x];
        __syncthreads();
        const int lim = (int)(hi - t0 < 256 ? hi - t0 : 256);
        for (int q = 0; q < lim; ++q) {
            const long long qb = tile[q];
            const int64_t b = t0 + q;
            r += (qb < qa || (qb == qa && b < a)) ? 1 : 0;
        }
    }
    if (a < count && r) atomicAdd(&rank[a], r);
}

__global__ void k_rank_scatter(const int32_t* rank, int64_t count, int64_t k, int64_t* sel);

void launch_rank_q(const unsigned long long* qi, int64_t count, int32_t* rank, hipStream_t s) {
    const unsigned gx = (unsigned)((count + 255) / 256);
    int64_t splits = (count + 63) / 64;  // about 64 comparisons per thread, at most 64 splits
    if (splits > 64) splits = 64;
    const int64_t span = (count + splits - 1) / splits;
    hipLaunchKernelGGL(k_rank_q, dim3(gx, (unsigned)splits), dim3(256), 0, s, qi, count, span, rank);
}

void launch_topk_q(const unsigned long long* qi, int64_t count, int64_t k, int32_t* rank, int64_t* sel,
                   hipStream_t s) {
    launch_rank_q(qi, count, rank, s);
    hipLaunchKernelGGL(k_rank_scatter, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, rank, count, k, sel);
}

// Each work group recomputes q_base = s^T h (n terms, int64, exact) and
// converts its 256 seeds; the rank counters of the following top-k start at 0.
__global__ __launch_bounds__(256) void k_seed_delta_i8(const uint8_t* x, const int32_t* h, int64_t n,
                                                       const unsigned long long* qi, int64_t count, double* delta,
                                                       int32_t* rank) {
    __shared__ long long part[4];
    long long acc = 0;
    for (int64_t v = threadIdx.x; v < n; v += blockDim.x) acc += x[v] ? -(long long)h[v] : (long long)h[v];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    const long long qb = part[0] + part[1] + part[2] + part[3];
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b < count) {
        delta[b] = (double)(qb - (long long)qi[b]) * 0.25;
        rank[b] = 0;
    }
}

void launch_seed_delta_i8(const uint8_t* x, const int32_t* h, int64_t n, const unsigned long long* qi,
                          int64_t count, double* delta, int32_t* rank, hipStream_t s) {
    hipLaunchKernelGGL(k_seed_delta_i8, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, x, h, n, qi, count,
                       delta, rank);
}

// Cut difference of flipping the set F from (x, h):
//   sum_{k in F} s_k h_k - 2 sum_{k<l in F} w_kl s_k s_l.
template <typename WT, typename FT>
__device__ double seed_lowrank_one(const GView<WT>& g, const uint8_t* x, const FT* h, int64_t bs, int64_t b) {
    using AT = typename AccOf<FT>::T;
    const int64_t n = g.n;
    const uint64_t i = (uint64_t)(b % bs);
    const uint64_t gray = i ^ (i >> 1);
    const int64_t offset = (b / bs) << 6;
    AT acc = 0;
    if (n >= 64) {
        // the 64 positions are distinct: walk the set bits of the code
        for (uint64_t rest = gray; rest; rest &= rest - 1) {
            const int64_t k = seed_pos(n, offset, __builtin_ctzll(rest));
            const int sk = x[k] ? -1 : 1;
            acc += (AT)sk * (AT)h[k];
            for (uint64_t r2 = rest & (rest - 1); r2; r2 &= r2 - 1) {
                const int64_t l = seed_pos(n, offset, __builtin_ctzll(r2));
                const int sl = x[l] ? -1 : 1;
                acc -= (AT)(2 * sk * sl) * (AT)weight_of(g, k, l);
            }
        }
        return (double)acc;
    }
    // n < 64: positions repeat, and a node flipped twice is not flipped; the
    // set is a mask over the n nodes
    uint64_t set = 0;
    for (int j = 0; j < 64 && j < n; ++j)
        if ((gray >> j) & 1u) set ^= 1ull << seed_pos(n, offset, j);
    for (uint64_t rest = set; rest; rest &= rest - 1) {
        const int64_t k = __builtin_ctzll(rest);
        const int sk = x[k] ? -1 : 1;
        acc += (AT)sk * (AT)h[k];
        for (uint64_t r2 = rest & (rest - 1); r2; r2 &= r2 - 1) {
            const int64_t l = __builtin_ctzll(r2);
            const int sl = x[l] ? -1 : 1;
            acc -= (AT)(2 * sk * sl) * (AT)weight_of(g, k, l);
        }
    }
    return (double)acc;
}

template <typename WT, typename FT>
__global__ __launch_bounds__(256) void k_seed_lowrank(GView<WT> g, const uint8_t* x, const FT* h, int64_t bs,
                                                      int64_t count, double* delta) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b < count) delta[b] = seed_lowrank_one<WT, FT>(g, x, h, bs, b);
}

template <typename WT>
static void lowrank_impl(const Graph& g, const uint8_t* x, const void* h, int64_t bs, int64_t count, double* delta,
                         hipStream_t s) {
    using FT = typename FieldOf<WT>::T;
    hipLaunchKernelGGL((k_seed_lowrank<WT, FT>), dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s,
                       view<WT>(g), x, (const FT*)h, bs, count, delta);
}

void launch_seed_lowrank(const Graph& g, const uint8_t* x, const void* h, int64_t bs, int64_t count, double* delta,
                         hipStream_t s) {
    PQI_DISPATCH(g, lowrank_impl, g, x, h, bs, count, delta, s);
}

__global__ void k_q_to_delta(const double* q, const double* q_base, int64_t count, double* delta) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b < count) delta[b] = (q_base[0] - q[b]) * 0.25;
}

void launch_q_to_delta(const double* q, const double* q_base, int64_t count, double* delta, hipStream_t s) {
    hipLaunchKernelGGL(k_q_to_delta, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, q, q_base, count, delta);
}

__global__ void k_gather(const double* v, const int64_t* sel, int64_t k, double* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < k) out[i] = v[sel[i]];
}

void launch_gather(const double* v, const int64_t* sel, int64_t k, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_gather, dim3((unsigned)((k + 255) / 256)), dim3(256), 0, s, v, sel, k, out);
}

// Stable top-k by rank counting: rank(a) = #{b : v_b > v_a or (v_b == v_a and b < a)}.
// Stage 1 splits the comparisons of each candidate over gridDim.y work groups
// (integer atomics, so the count is exact); stage 2 scatters the first k.
__device__ void rank_body(const double* v, int64_t count, int64_t span, int32_t* rank, int64_t bx, int64_t by) {
    __shared__ double tile[256];
    const int64_t a = bx * blockDim.x + threadIdx.x;
    const double va = a < count ? v[a] : 0.0;
    const int64_t lo = by * span;
    const int64_t hi = lo + span < count ? lo + span : count;
    int32_t r = 0;
    for (int64_t t0 = lo; t0 < hi; t0 += 256) {
        __syncthreads();
        if (t0 + threadIdx.x < hi) tile[threadIdx.x] = v[t0 + threadIdx.x];
        __syncthreads();
        const int lim = (int)(hi - t0 < 256 ? hi - t0 : 256);
        for (int q = 0; q < lim; ++q) {
            const double vb = tile[q];
            const int64_t b = t0 + q;
            r += (vb > va || (vb == va && b < a)) ? 1 : 0;
        }
    }
    if (a < count && r) atomicAdd(&rank[a], r);
}

__global__ __launch_bounds__(256) void k_rank(const double* v, int64_t count, int64_t span, int32_t* rank) {
    rank_body(v, count, span, rank, blockIdx.x, blockIdx.y);
}

__global__ void k_rank_scatter(const int32_t* rank, int64_t count, int64_t k, int64_t* sel) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a < count && rank[a] < k) sel[rank[a]] = a;
}

__global__ void k_zero_i32(int32_t* p, int64_t count) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0;
}

size_t topk_scratch_bytes(int64_t count) { return (size_t)count * 4; }

__global__ void k_rank_scatter_gather(const int32_t* rank, const double* v, int64_t count, int64_t k, int64_t* sel,
                                      double* seld) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a < count && rank[a] < k) {
        sel[rank[a]] = a;
        seld[rank[a]] = v[a];
    }
}

void launch_topk_gather(const double* v, int64_t count, int64_t k, int64_t* sel, double* seld, int32_t* rank,
                        hipStream_t s) {
    const unsigned gx = (unsigned)((count + 255) / 256);
    int64_t splits = (count + 63) / 64;  // about 64 comparisons per thread, at most 64 splits
    if (splits > 64) splits = 64;
    const int64_t span = (count + splits - 1) / splits;
    hipLaunchKernelGGL(k_rank, dim3(gx, (unsigned)splits), dim3(256), 0, s, v, count, span, rank);
    hipLaunchKernelGGL(k_rank_scatter_gather, dim3(gx), dim3(256), 0, s, rank, v, count, k, sel, seld);
}

void launch_topk(const double* v, int64_t count, int64_t k, int64_t* sel, void* scratch, hipStream_t s) {
    int32_t* rank = (int32_t*)scratch;
    const unsigned gx = (unsigned)((count + 255) / 256);
    int64_t splits = (count + 63) / 64;  // about 64 comparisons per thread, at most 64 splits
    if (splits > 64) splits = 64;
    const int64_t span = (count + splits - 1) / splits;
    hipLaunchKernelGGL(k_zero_i32, dim3(gx), dim3(256), 0, s, rank, count);
    hipLaunchKernelGGL(k_rank, dim3(gx, (unsigned)splits), dim3(256), 0, s, v, count, span, rank);
    hipLaunchKernelGGL(k_rank_scatter, dim3(gx), dim3(256), 0, s, rank, count, k, sel);
}

// ---------------------------------------------------------------------------
// Gray-code walkers (pyqrackising/spin_glass_solver.py:126-153).
//
// One wavefront per walker.  Iteration i flips, for block b = 0..nb-1 in order,
// the node k = ctz(i+1) + 64 b and keeps it iff its gain is positive.  All nb
// candidate gains of an iteration are evaluated at once, one block per lane;
// a ballot finds the first positive one, which is applied, and the lanes after
// it re-evaluate.  This is the reference's sequential order exactly.
// The walker's state lives in LDS when it fits, else in its global slot.
// ---------------------------------------------------------------------------

// Skipping rule of the Gray walkers.  Iteration i (i + 1 = m) evaluates the
// flips at position ctz(m); `clean` has bit p set when every flip at position
// p has been evaluated and rejected since the walker last changed, so those
// iterations repeat exactly and are skipped.  Returns the smallest m >= v whose
// position is not clean, or 0 when there is none up to T (pmax = floor(log2 T)).
// Positions below p0 = ctz(~clean) are all clean, so the answer is m0, the
// smallest odd multiple of 2^p0 that is >= v, unless the one multiple of 2^p0
// below it, t = m0 - 2^p0, is >= v and sits at a position that is not clean
// (t is an even multiple of 2^p0, so its position is above p0).
__device__ __forceinline__ uint64_t next_unclean(uint64_t v, uint64_t clean, int pmax, uint64_t T) {
    const uint64_t nc = ~clean;
    if (!nc) return 0;
    const int p0 = __builtin_ctzll(nc);
    if (p0 > pmax) return 0;
    const uint64_t step = 1ull << p0;
    uint64_t m0 = (v + step - 1) & ~(step - 1);
    if (!(m0 & step)) m0 += step;
    uint64_t best = m0;
    const uint64_t t = m0 - step;
    if (t >= v && t != 0 && ((nc >> __builtin_ctzll(t)) & 1u)) best = t;
    return best <= T ? best : 0;
}

template <typename WT, typename FT>
__device__ void gray_walk_body(const GView<WT>& g, const uint8_t* x0, const FT* h0, const int64_t* sel,
                               const double* sel_delta, int64_t bs, int64_t T, uint8_t* wx, FT* wh, double* we,
                               int64_t* wevals, int use_lds, unsigned char* lds, int64_t w) {
    const int lane = threadIdx.x;
    const int64_t n = g.n;
    const int64_t nb = (n + 63) >> 6;
    uint8_t* gx = wx + w * n;
    FT* gh = wh + w * n;
    FT* h = use_lds ? (FT*)lds : gh;
    uint8_t* x = use_lds ? (uint8_t*)(lds + ((n * sizeof(FT) + 15) & ~(size_t)15)) : gx;

    for (int64_t v = lane; v < n; v += 64) {
        h[v] = h0[v];
        x[v] = x0[v];
    }
    __syncthreads();

    // seed mutation
    const int64_t seed = sel[w];
    const uint64_t si = (uint64_t)(seed % bs);
    const uint64_t gray = si ^ (si >> 1);
    const int64_t offset = (seed / bs) << 6;
    for (int j = 0; j < 64 && j < n; ++j) {
        if (!((gray >> j) & 1u)) continue;
        const int64_t k = seed_pos(n, offset, j);
        const int s_old = x[k] ? -1 : 1;
        __syncthreads();
        flip_update(g, x, h, k, s_old, lane, 64);
        __syncthreads();
    }

    double energy = sel_delta[w];
    int64_t evals = 0;
    // clean bit p: every flip at positions p + 64 b has been evaluated and
    // rejected since the walker last changed.  A rejected flip is reverted, so
    // until the next accepted flip those iterations repeat exactly and are
    // skipped (their gain is 0).  pmax: largest p with an iteration below T.
    uint64_t clean = 0;
    const int pmax = 63 - __builtin_clzll((unsigned long long)(T > 0 ? T : 1));
    for (int64_t ci = 0; ci < T;) {
        const int p = __builtin_ctzll((unsigned long long)(ci + 1));
        if ((clean >> p) & 1u) {
            const uint64_t next = next_unclean((uint64_t)ci + 1, clean, pmax, (uint64_t)T);
            if (!next) break;
            ci = (int64_t)next - 1;
            continue;
        }
        bool changed = false;
        double gain_i = 0.0;
        for (int64_t b0 = 0; b0 < nb; b0 += 64) {
            const int64_t b = b0 + lane;
            const int64_t k = (int64_t)p + 64 * b;
            const bool valid = (b < nb) && (k < n);
            FT d = valid ? gain_of(x, h, k) : (FT)0;
            evals += __popcll(__ballot(valid));
            uint64_t mask = __ballot(valid && d > 0);
            while (mask) {
                const int bl = __builtin_ctzll(mask);
                const int64_t kk = (int64_t)p + 64 * (b0 + bl);
                const FT dd = __shfl(d, bl, 64);
                gain_i += (double)dd;
                changed = true;
                const int s_old = x[kk] ? -1 : 1;
                __syncthreads();
                flip_update(g, x, h, kk, s_old, lane, 64);
                __syncthreads();
                d = valid ? gain_of(x, h, k) : (FT)0;
                mask = __ballot(valid && d > 0 && lane > bl);
            }
        }
        energy += gain_i;
        clean = changed ? 0ull : (clean | (1ull << p));
        ++ci;
    }
    __syncthreads();
    if (use_lds) {
        for (int64_t v = lane; v < n; v += 64) {
            gh[v] = h[v];
            gx[v] = x[v];
        }
    }
    if (lane == 0) {
        we[w] = energy;
        wevals[w] = evals;
    }
}

// Register-resident Gray walker for int8 weights with a fixed-stride neighbour
// table (n <= 2048, rows of at most NB entries).  Node v lives in register
// v & 31 of lane (v >> 6) + 32 ((v >> 5) & 1): lanes 0-31 hold the low halves
// of the 64-node blocks, lanes 32-63 the high halves, so the candidates of an
// iteration at position p (nodes p + 64 b) are register p & 31 of the lanes of
// one half, in block order: one uniform v_movrel read, no memory access.
// word = (Delta << 1) | x.  A flip updates the neighbours one at a time (the
// row comes from the table, staged in LDS when it fits).  Same iteration
// order, acceptance rule, skipping rule, energy and evaluation count as
// gray_walk_body.  The host enables it only when every row strength is below
// 2^29, so Delta << 1 fits.
__device__ __forceinline__ int walk_lane_of(int v) { return (v >> 6) | (((v >> 5) & 1) << 5); }

template <int NB>
__device__ void gray_walk_reg_body(const GView<int8_t>& g, const uint8_t* x0, const int32_t* h0, const int64_t* sel,
                                   const double* sel_delta, int64_t bs, int64_t T, uint8_t* wx, int32_t* wh,
                                   double* we, int64_t* wevals, unsigned char* lds, size_t lds_bytes, int64_t w) {
    const int lane = threadIdx.x;
    const int n = (int)g.n;
    const int nb = (n + 63) >> 6;
    const int stride = (int)g.kstride;
    const int base = ((lane & 31) << 6) | ((lane >> 5) << 5);  // node of register 0 in this lane
    // a lane's registers hold 32 consecutive nodes: whole blocks are read in
    // 16-byte pieces when the state is 16-byte aligned
    const bool full = (lane & 31) < nb && base + 32 <= n;
    int32_t wr[32];
    if (full && (((uintptr_t)x0 | (uintptr_t)h0) & 15) == 0) {
        const uint4* xs = (const uint4*)(x0 + base);
        const int4* hs = (const int4*)(h0 + base);
        const uint4 xa = xs[0], xb = xs[1];
        int4 hq[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) hq[q] = hs[q];
        const uint32_t xw[8] = {xa.x, xa.y, xa.z, xa.w, xb.x, xb.y, xb.z, xb.w};
#pragma unroll
        for (int r = 0; r < 32; ++r) {
            const int32_t xv = (int32_t)((xw[r >> 2] >> (8 * (r & 3))) & 0xFFu);
            const int4 hp = hq[r >> 2];
            const int32_t hv = (r & 3) == 0 ? hp.x : (r & 3) == 1 ? hp.y : (r & 3) == 2 ? hp.z : hp.w;
            wr[r] = (int32_t)((uint32_t)(xv ? -hv : hv) << 1) | xv;
        }
    } else {
#pragma unroll
        for (int r = 0; r < 32; ++r) {
            const int v = base + r;
            int32_t word = 0;
            if ((lane & 31) < nb && v < n) {
                const int32_t xv = x0[v];
                word = (int32_t)((uint32_t)(xv ? -h0[v] : h0[v]) << 1) | xv;
            }
            wr[r] = word;
        }
    }
    const int quads = (n * stride + 3) >> 2;
    const bool staged = (size_t)quads * 16 <= lds_bytes;
    if (staged) {
        const int4* src = (const int4*)g.kadj;
        int4* dst = (int4*)lds;
#pragma unroll 12  // 12 loads in flight per lane: a 24 KiB table in two rounds
        for (int q = lane; q < quads; q += 64) dst[q] = src[q];
    }
    __syncthreads();
    // flip node K (register K & 31 of lane walk_lane_of(K)); SK2 = 2 s_K before the flip
    // apply the flip of node K whose neighbour row entry j is ROW(j) (uniform);
    // SK2 = 2 s_K before the flip
#define WFLIP_ROW(K, SK2, ROW)                                                                               \
    {                                                                                                      \
        const int fk = (K), fsk2 = (SK2);                                                                  \
        _Pragma("unroll") for (int j = 0; j < NB; ++j) {                                                   \
            if (j >= stride) break;                                                                        \
            const int32_t a = ROW(j); /* a zero padding entry (w = 0) changes nothing */                   \
            const int u = a & 0xFFFFFF;                                                                    \
            const int slot = u & 31;                                                                       \
            const int sk2w = fsk2 * (a >> 24);                                                             \
            const int32_t kv = wr[slot];                                                                   \
            const int32_t sm = __builtin_amdgcn_sbfe(kv, 0u, 1u);                                          \
            const int32_t nv = (int32_t)((uint32_t)kv - ((uint32_t)((sk2w ^ sm) - sm) << 1));               \
            wr[slot] = lane == walk_lane_of(u) ? nv : kv;                                                  \
        }                                                                                                  \
        const int kslot = fk & 31;                                                                         \
        const int32_t kv = wr[kslot];                                                                      \
        const int32_t nv = (int32_t)(((uint32_t)(-(kv >> 1)) << 1) | ((kv & 1) ^ 1));                      \
        wr[kslot] = lane == walk_lane_of(fk) ? nv : kv;                                                    \
    }
    // flip node K, its row read by lanes 0..stride-1
#define WFLIP(K, SK2)                                                                                        \
    {                                                                                                      \
        const int col = lane < stride ? lane : 0;                                                          \
        const int32_t a0 = staged ? ((const int32_t*)lds)[(K) * stride + col]                              \
                                  : g.kadj[(int64_t)(K) * stride + col];                                   \
        auto row_of = [&](int j) { return __builtin_amdgcn_readlane(a0, j); };                             \
        WFLIP_ROW(K, SK2, row_of)                                                                          \
    }

    // seed mutation: lane j loads the neighbour row of the j-th mutated node
    // up front, so the sequential flips read their rows from registers
    const int64_t seed = sel[w];
    const uint64_t si = (uint64_t)(seed % bs);
    const uint64_t gray = si ^ (si >> 1);
    const int64_t offset = (seed / bs) << 6;
    {
        const int kj = lane < n ? (int)seed_pos(n, offset, lane) : 0;
        int32_t rowj[NB];
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const int col = j < stride ? j : 0;
            rowj[j] = staged ? ((const int32_t*)lds)[kj * stride + col] : g.kadj[(int64_t)kj * stride + col];
        }
        for (int j = 0; j < 64 && j < n; ++j) {
            if (!((gray >> j) & 1u)) continue;
            const int k = (int)seed_pos(n, offset, j);
            const int32_t wk = __builtin_amdgcn_readlane(wr[k & 31], walk_lane_of(k));
            auto row_of = [&](int i) { return __builtin_amdgcn_readlane(rowj[i], j); };
            WFLIP_ROW(k, (wk & 1) ? -2 : 2, row_of)
        }
    }

    double energy = sel_delta[w];
    int64_t evals = 0;
    // clean bit p: as in gray_walk_body
    uint64_t clean = 0;
    const int pmax = 63 - __builtin_clzll((unsigned long long)(T > 0 ? T : 1));
    for (int64_t ci = 0; ci < T;) {
        const int p = __builtin_ctzll((unsigned long long)(ci + 1));
        if ((clean >> p) & 1u) {
            const uint64_t next = next_unclean((uint64_t)ci + 1, clean, pmax, (uint64_t)T);
            if (!next) break;
            ci = (int64_t)next - 1;
            continue;
        }
        const int ps = p & 31;
        const bool valid = (lane >> 5) == (p >> 5) && (lane & 31) < nb && base + ps < n;
        int32_t word = wr[ps];
        evals += __popcll(__ballot(valid));
        uint64_t mask = __ballot(valid && (word >> 1) > 0);
        bool changed = false;
        double gain_i = 0.0;
        while (mask) {
            const int bl = __builtin_ctzll(mask);
            const int32_t wk = __builtin_amdgcn_readlane(word, bl);
            gain_i += (double)(wk >> 1);
            changed = true;
            WFLIP(p + 64 * (bl & 31), (wk & 1) ? -2 : 2);
            word = wr[ps];
            mask = __ballot(valid && (word >> 1) > 0 && lane > bl);
        }
        energy += gain_i;
        clean = changed ? 0ull : (clean | (1ull << p));
        ++ci;
    }
#undef WFLIP
#undef WFLIP_ROW
    uint8_t* gx = wx + w * n;
    int32_t* gh = wh + w * n;
    if (full && (((uintptr_t)gx | (uintptr_t)gh) & 15) == 0) {
        int32_t hv[32];
        uint32_t xw[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int r = 0; r < 32; ++r) {
            const int32_t word = wr[r];
            const int32_t d = word >> 1;
            hv[r] = (word & 1) ? -d : d;
            xw[r >> 2] |= (uint32_t)(word & 1) << (8 * (r & 3));
        }
        int4* hs = (int4*)(gh + base);
#pragma unroll
        for (int q = 0; q < 8; ++q) hs[q] = make_int4(hv[4 * q], hv[4 * q + 1], hv[4 * q + 2], hv[4 * q + 3]);
        uint4* xs = (uint4*)(gx + base);
        xs[0] = make_uint4(xw[0], xw[1], xw[2], xw[3]);
        xs[1] = make_uint4(xw[4], xw[5], xw[6], xw[7]);
    } else {
#pragma unroll
        for (int r = 0; r < 32; ++r) {
            const int v = base + r;
            if ((lane & 31) < nb && v < n) {
                const int32_t word = wr[r];
                const int32_t d = word >> 1;
                gh[v] = (word & 1) ? -d : d;
                gx[v] = (uint8_t)(word & 1);
            }
        }
    }
    if (lane == 0) {
        we[w] = energy;
        wevals[w] = evals;
    }
}

template <int NB>
__global__ __launch_bounds__(64) void k_gray_walk_reg(GView<int8_t> g, const uint8_t* x0, const int32_t* h0,
                                                     const int64_t* sel, const double* sel_delta, int64_t bs,
                                                     int64_t T, uint8_t* wx, int32_t* wh, double* we,
                                                     int64_t* wevals, size_t lds_bytes, const int32_t* gate) {
    extern __shared__ __align__(16) unsigned char lds[];
    if (gate && *gate) return;
    gray_walk_reg_body<NB>(g, x0, h0, sel, sel_delta, bs, T, wx, wh, we, wevals, lds, lds_bytes, blockIdx.x);
}

template <typename WT, typename FT>
__global__ __launch_bounds__(64) void k_gray_walk(GView<WT> g, const uint8_t* x0, const FT* h0, const int64_t* sel,
                                                 const double* sel_delta, int64_t bs, int64_t T, uint8_t* wx,
                                                 FT* wh, double* we, int64_t* wevals, int use_lds,
                                                 const int32_t* gate) {
    extern __shared__ __align__(16) unsigned char lds[];
    if (gate && *gate) return;
    gray_walk_body<WT, FT>(g, x0, h0, sel, sel_delta, bs, T, wx, wh, we, wevals, use_lds, lds, blockIdx.x);
}

template <int NB>
static void launch_walk_reg(const Graph& g, const uint8_t* x, const void* h, const int64_t* sel,
                            const double* sel_delta, int64_t bs, int64_t walkers, int64_t T, uint8_t* wx, void* wh,
                            double* we, int64_t* wevals, size_t table, hipStream_t s, const int32_t* gate) {
    static const size_t limit = lds_limit((const void*)k_gray_walk_reg<NB>);
    const size_t lds = table <= limit ? table : 0;
    hipLaunchKernelGGL(k_gray_walk_reg<NB>, dim3((unsigned)walkers), dim3(64), lds, s, view<int8_t>(g), x,
                       (const int32_t*)h, sel, sel_delta, bs, T, wx, (int32_t*)wh, we, wevals, lds, gate);
}

template <typename WT>
static void walk_impl(const Graph& g, const uint8_t* x, const void* h, const int64_t* sel, const double* sel_delta,
                      int64_t bs, int64_t walkers, int64_t T, uint8_t* wx, void* wh, double* we, int64_t* wevals,
                      hipStream_t s, const int32_t* gate) {
    using FT = typename FieldOf<WT>::T;
    if constexpr (sizeof(WT) == 1) {
        size_t table = 0;
        if (g.n <= 2048 && reg_descent_plan(g, &table)) {
            switch (reg_nb(g.key_adj_stride)) {
                case 4: launch_walk_reg<4>(g, x, h, sel, sel_delta, bs, walkers, T, wx, wh, we, wevals, table, s,
                                           gate); break;
                case 8: launch_walk_reg<8>(g, x, h, sel, sel_delta, bs, walkers, T, wx, wh, we, wevals, table, s,
                                           gate); break;
                default: launch_walk_reg<16>(g, x, h, sel, sel_delta, bs, walkers, T, wx, wh, we, wevals, table, s,
                                             gate); break;
            }
            return;
        }
    }
    size_t lds = ((g.n * sizeof(FT) + 15) & ~(size_t)15) + (size_t)g.n;
    static const size_t limit = lds_limit((const void*)k_gray_walk<WT, FT>);
    int use_lds = lds <= limit ? 1 : 0;
    hipLaunchKernelGGL((k_gray_walk<WT, FT>), dim3((unsigned)walkers), dim3(64), use_lds ? lds : 0, s, view<WT>(g),
                       x, (const FT*)h, sel, sel_delta, bs, T, wx, (FT*)wh, we, wevals, use_lds, gate);
}

void launch_gray_walk(const Graph& g, const uint8_t* x, const void* h, const int64_t* sel, const double* sel_delta,
                      int64_t block_size, int64_t walkers, int64_t thread_iterations, uint8_t* wx, void* wh,
                      double* we, int64_t* wevals, hipStream_t s, const int32_t* gate) {
    PQI_DISPATCH(g, walk_impl, g, x, h, sel, sel_delta, block_size, walkers, thread_iterations, wx, wh, we, wevals, s,
                 gate);
}

// ---------------------------------------------------------------------------
// Device-resident control of the single-instance search loop
// (pyqrackising/spin_glass_solver.py:603-717).  Each control kernel reads the
// previous phase's PhaseOut, applies the accepted move to (x, h) with the
// whole work group, and updates the counters and max_energy in the order of
// the reference loop.  Control values are read by every thread before thread
// 0 writes any of them, so branches stay uniform.
// ---------------------------------------------------------------------------

__device__ __forceinline__ void ls_mark(LsCtl* c, int phase) {
    const uint64_t now = wall_clock64();
    if (c->phase >= 0) c->ticks[c->phase] += now - c->mark;
    c->mark = now;
    c->phase = phase;
}

// The reset of an iteration.  When the search has stopped (done, or the round
// bound reached outside a continued descent), `restart` stays set instead, so
// every phase of a further iteration already queued is skipped.
__device__ __forceinline__ void ls_begin_body(LsCtl* c) {
    if (ls_stopped(c)) {
        c->restart = 1;
        return;
    }
    if (!c->again) c->rounds++;
    c->again = 0;
    c->restart = 0;
    c->skip_seed = 1;
    c->seeds_ran = 0;
    c->walk_ran = 0;
    ls_mark(c, 0);
}
__global__ void k_ls_begin(LsCtl* c) { ls_begin_body(c); }

// Control steps as work-group bodies: thread 0 writes LsCtl; the bodies that
// apply a move use the whole work group.  The kernels below run one body, or
// several in the iteration's order (fused, one kernel boundary instead of
// several); between bodies the group re-reads `restart` after a barrier, so
// every thread takes the same branch.
__device__ __forceinline__ bool ls_restarted(const LsCtl* c) {
    __shared__ int r;
    __syncthreads();
    if (threadIdx.x == 0) r = c->restart;
    __syncthreads();
    return r != 0;
}

__device__ __forceinline__ void ls_after_descent_body(LsCtl* c, const PhaseOut* o, int64_t cap) {
    c->single_evals += o->evals;
    c->descent_evals += o->evals;
    c->single_accepts += o->accepts;
    c->max_energy += c->scale * o->gain_sum;
    if (o->accepts == cap) {
        // the bound was hit: continue the descent in the next iteration (same round)
        c->again = 1;
        c->restart = 1;
    } else {
        ls_mark(c, 1);
    }
}

__global__ void k_ls_after_descent(LsCtl* c, const PhaseOut* o, int64_t cap) {
    if (c->restart) return;  // the search stopped: the descent was gated off
    ls_after_descent_body(c, o, cap);
}

template <typename FT, typename WT>
__device__ __forceinline__ void copy_state(uint8_t* __restrict__ xd, FT* __restrict__ hd,
                                           const uint8_t* __restrict__ xs, const FT* __restrict__ hs, int64_t n) {
#pragma unroll 4
    for (int64_t v = threadIdx.x; v < n; v += blockDim.x) {
        xd[v] = xs[v];
        hd[v] = hs[v];
    }
}

template <typename WT, typename FT>
__device__ void ls_after_double_body(const GView<WT>& g, LsCtl* c, const PhaseOut* o, uint8_t* x, FT* h,
                                     int64_t evals) {
    const double gain = o->gain;
    const int64_t k = o->k, l = o->l;
    __syncthreads();
    if (gain > 0.0) flip_body<WT, FT>(g, x, h, k, l);
    if (threadIdx.x == 0) {
        c->double_evals += evals;
        if (gain > 0.0) {
            c->max_energy += c->scale * gain;
            c->double_accepts++;
            c->restart = 1;
        } else {
            c->seeds_ran = 1;
            ls_mark(c, 2);
        }
    }
}

template <typename WT, typename FT>
__global__ __launch_bounds__(1024) void k_ls_after_double(GView<WT> g, LsCtl* c, const PhaseOut* o, uint8_t* x,
                                                         FT* h, int64_t evals) {
    if (c->restart) return;
    ls_after_double_body<WT, FT>(g, c, o, x, h, evals);
}

// After the descent: its bookkeeping, the reference-mode double flip
// (spin_glass_solver.py:55-92) and its decision, in work group 0.  With seed
// rows (int8 MFMA seeds), the other work groups write the Gray-seed rows of the
// state meanwhile.  The rows are used only when the double flip is rejected,
// which leaves the state as the descent left it; when it is accepted (or the
// descent restarts the iteration) they are not used, and these groups may skip
// them as soon as they see `restart` set.
template <typename WT, typename FT>
__global__ __launch_bounds__(1024) void k_ls_descent_double(GView<WT> g, LsCtl* c, const PhaseOut* dout,
                                                           int64_t cap, uint8_t* x, FT* h, int64_t thread_count,
                                                           PhaseOut* dout2, int64_t evals, SeedRows sr) {
    if (blockIdx.x > 0) {
        if (*(volatile int32_t*)&c->restart) return;
        seed_bits_body((int64_t)(blockIdx.x - 1) * blockDim.x + threadIdx.x, x, g.n, sr.wpr, sr.bs, sr.count, sr.bits,
                       sr.qzero, sr.rzero);
        return;
    }
    if (ls_restarted(c)) return;  // the search stopped: the descent was gated off
    if (threadIdx.x == 0) ls_after_descent_body(c, dout, cap);
    if (ls_restarted(c)) return;
    double_ref_body<WT, FT>(g, x, h, thread_count, dout2);
    __syncthreads();
    ls_after_double_body<WT, FT>(g, c, dout2, x, h, evals);
}

__device__ __forceinline__ void ls_after_seeds_body(LsCtl* c, const double* seld, int64_t count) {
    c->seed_evals += count;
    c->top = seld[0];
    c->skip_seed = c->top > 0.0 ? 0 : 1;
}

// The Gray-seed decision and, when the best seed improves, its mutation applied
// to (x, h) in place: the same flips, in the same order and with the same
// field arithmetic, as the one-walker, zero-iteration walk followed by the
// copy of its state (gray_mutation, maxcut_tfim_util.py:740-749).  With rank
// (the seeds' stable ranks), the top G are first placed into sel here
// (k_rank_scatter's step).
template <typename WT, typename FT>
__global__ __launch_bounds__(1024) void k_ls_seed_take(GView<WT> g, LsCtl* c, double* seld, int64_t count,
                                                      int64_t* sel, int64_t bs, uint8_t* x, FT* h,
                                                      const unsigned long long* qi, int64_t G, const int32_t* rank) {
    __shared__ int take;
    __shared__ long long part[16];
    if (qi) {
        // int8 MFMA seeds: the selected seeds' cut differences (q_base - q_b) / 4,
        // q_base = s^T h of the state (k_seed_delta_i8's arithmetic)
        if (ls_restarted(c)) return;
        if (rank) {
            for (int64_t a = threadIdx.x; a < count; a += blockDim.x) {
                const int32_t r = rank[a];
                if (r >= 0 && r < G) sel[r] = a;
            }
            __syncthreads();
        }
        long long acc = 0;
        for (int64_t v = threadIdx.x; v < g.n; v += blockDim.x) acc += x[v] ? -(long long)h[v] : (long long)h[v];
        acc = wave_sum(acc);
        if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
        __syncthreads();
        long long qb = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) qb += part[w];
        for (int64_t r = threadIdx.x; r < G; r += blockDim.x) seld[r] = (double)(qb - (long long)qi[sel[r]]) * 0.25;
        __syncthreads();
    }
    if (threadIdx.x == 0 && !c->restart) ls_after_seeds_body(c, seld, count);
    if (ls_restarted(c)) return;
    if (threadIdx.x == 0) take = !c->skip_seed;
    __syncthreads();
    if (take) {
        const int64_t n = g.n;
        const int64_t seed = sel[0];
        const uint64_t si = (uint64_t)(seed % bs);
        const uint64_t gray = si ^ (si >> 1);
        const int64_t offset = (seed / bs) << 6;
        for (int j = 0; j < 64 && j < n; ++j) {
            if (!((gray >> j) & 1u)) continue;
            flip_body<WT, FT>(g, x, h, seed_pos(n, offset, j), -1);
            __syncthreads();
        }
    }
    if (threadIdx.x == 0) {
        if (take) {
            c->max_energy += c->scale * c->top;
            c->seed_accepts++;
            c->restart = 1;
        } else {
            c->walk_ran = 1;
            ls_mark(c, 3);
        }
    }
}

template <typename WT, typename FT>
__device__ void ls_after_walk_body(const GView<WT>& g, LsCtl* c, const double* we, const int64_t* wev,
                                   int64_t walkers, int64_t gray_evals, const uint8_t* wx, const FT* wh, uint8_t* x,
                                   FT* h, uint8_t* rx, FT* rh) {
    __shared__ double sv[16];
    __shared__ int64_t sk[16];
    __shared__ int64_t ssum[16];
    // first maximum over the walkers (value, then lowest walker) and the
    // number of computed evaluations
    double bv = -INFINITY;
    int64_t bw = LLONG_MAX, computed = 0;
    for (int64_t w = threadIdx.x; w < walkers; w += blockDim.x) {
        if (better(we[w], w, bv, bw)) { bv = we[w]; bw = w; }
        computed += wev[w];
    }
    computed = wave_sum(computed);
    if ((threadIdx.x & 63) == 0) ssum[threadIdx.x >> 6] = computed;
    block_argmax(bv, bw, sv, sk);  // synchronises
    const double walk = bv;
    const int64_t n = g.n;
    if (walk > 0.0) copy_state<FT, WT>(x, h, wx + bw * n, wh + bw * n, n);
    else copy_state<FT, WT>(rx, rh, wx + bw * n, wh + bw * n, n);
    if (threadIdx.x == 0) {
        computed = 0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) computed += ssum[i];
        c->gray_evals_computed += computed;
        c->gray_evals += gray_evals;
        if (walk > 0.0) {
            c->max_energy += c->scale * walk;
            c->walk_accepts++;
            c->restart = 1;
        } else {
            c->reheat = walk;
            ls_mark(c, 4);
        }
    }
}

template <typename WT, typename FT>
__global__ __launch_bounds__(1024) void k_ls_after_walk(GView<WT> g, LsCtl* c, const double* we, const int64_t* wev,
                                                       int64_t walkers, int64_t gray_evals, const uint8_t* wx,
                                                       const FT* wh, uint8_t* x, FT* h, uint8_t* rx, FT* rh) {
    if (c->restart) return;
    ls_after_walk_body<WT, FT>(g, c, we, wev, walkers, gray_evals, wx, wh, x, h, rx, rh);
}

template <typename WT, typename FT>
__device__ void ls_reheat_single_body(const GView<WT>& g, LsCtl* c, const PhaseOut* o, uint8_t* x, FT* h,
                                      uint8_t* rx, FT* rh) {
    const double gain = o->gain, reheat = c->reheat;
    const int64_t k = o->k;
    __syncthreads();
    const bool take = gain + reheat > 0.0;
    if (take) {
        copy_state<FT, WT>(x, h, rx, rh, g.n);
        __syncthreads();
        flip_body<WT, FT>(g, x, h, k, -1);
    } else if (gain > 0.0) {
        flip_body<WT, FT>(g, rx, rh, k, -1);
    }
    if (threadIdx.x == 0) {
        c->single_evals += o->evals;
        if (take) {
            c->max_energy += c->scale * (reheat + gain);
            c->reheat_accepts++;
            c->restart = 1;
        } else if (gain > 0.0) {
            c->reheat = reheat + gain;
        }
    }
}

template <typename WT, typename FT>
__global__ __launch_bounds__(1024) void k_ls_reheat_single(GView<WT> g, LsCtl* c, const PhaseOut* o, uint8_t* x,
                                                          FT* h, uint8_t* rx, FT* rh) {
    if (c->restart) return;
    ls_reheat_single_body<WT, FT>(g, c, o, x, h, rx, rh);
}

template <typename WT, typename FT>
__device__ void ls_reheat_double_body(const GView<WT>& g, LsCtl* c, const PhaseOut* o, uint8_t* x, FT* h,
                                      uint8_t* rx, FT* rh, int64_t evals) {
    const double gain = o->gain, reheat = c->reheat;
    const int64_t k = o->k, l = o->l;
    __syncthreads();
    const bool take = gain + reheat > 0.0;
    if (take) {
        copy_state<FT, WT>(x, h, rx, rh, g.n);
        __syncthreads();
        flip_body<WT, FT>(g, x, h, k, l);
    }
    if (threadIdx.x == 0) {
        c->double_evals += evals;
        if (take) {
            c->max_energy += c->scale * (reheat + gain);
            c->reheat_accepts++;
            c->restart = 1;
        } else {
            c->done = 1;
            ls_mark(c, -1);
        }
    }
}

template <typename WT, typename FT>
__global__ __launch_bounds__(1024) void k_ls_reheat_double(GView<WT> g, LsCtl* c, const PhaseOut* o, uint8_t* x,
                                                          FT* h, uint8_t* rx, FT* rh, int64_t evals) {
    if (c->restart) return;
    ls_reheat_double_body<WT, FT>(g, c, o, x, h, rx, rh, evals);
}

// After the main walk, in one kernel: its decision, then the post-reheat
// phase (spin_glass_solver.py:670-717) with the best single flip and the
// reference-mode double flip of the reheat state and their decisions.
template <typename WT, typename FT>
__device__ void ls_walk_reheat_body(const GView<WT>& g, LsCtl* c, const double* we, const int64_t* wev,
                                    int64_t walkers, int64_t gray_evals, const uint8_t* wx, const FT* wh, uint8_t* x,
                                    FT* h, uint8_t* rx, FT* rh, PhaseOut* dout, int64_t thread_count,
                                    PhaseOut* dout2, int64_t dbl_evals) {
    if (ls_restarted(c)) return;
    ls_after_walk_body<WT, FT>(g, c, we, wev, walkers, gray_evals, wx, wh, x, h, rx, rh);
    if (ls_restarted(c)) return;
    descent_body<WT, FT>(g, rx, rh, 0, dout);  // best single flip of the reheat state, no move
    __syncthreads();
    ls_reheat_single_body<WT, FT>(g, c, dout, x, h, rx, rh);
    if (ls_restarted(c)) return;
    double_ref_body<WT, FT>(g, rx, rh, thread_count, dout2);
    __syncthreads();
    ls_reheat_double_body<WT, FT>(g, c, dout2, x, h, rx, rh, dbl_evals);
}

__device__ void ls_report_body(const uint8_t* x, const int32_t* h, int64_t n, int64_t w_up_i, LsCtl* c,
                               const FieldRepair* fix, LsReport* rep, uint8_t* x_out, unsigned long long* seq);

// With a report (int8 weights only), the iteration's report and the next
// iteration's control reset (k_ls_report_i8's body) follow in the same kernel,
// whether or not the iteration had already restarted.
template <typename WT, typename FT>
__global__ __launch_bounds__(1024) void k_ls_walk_reheat(GView<WT> g, LsCtl* c, const double* we, const int64_t* wev,
                                                        int64_t walkers, int64_t gray_evals, const uint8_t* wx,
                                                        const FT* wh, uint8_t* x, FT* h, uint8_t* rx, FT* rh,
                                                        PhaseOut* dout, int64_t thread_count, PhaseOut* dout2,
                                                        int64_t dbl_evals, LsReportTo rt) {
    ls_walk_reheat_body<WT, FT>(g, c, we, wev, walkers, gray_evals, wx, wh, x, h, rx, rh, dout, thread_count, dout2,
                                dbl_evals);
    if constexpr (std::is_same<FT, int32_t>::value) {
        if (rt.rep) {
            __syncthreads();
            ls_report_body(x, h, g.n, rt.w_up_i, c, rt.fix, rt.rep, rt.x_out, rt.seq);
        }
    }
}

void launch_ls_begin(LsCtl* c, hipStream_t s) { hipLaunchKernelGGL(k_ls_begin, dim3(1), dim3(1), 0, s, c); }

__global__ void k_ls_start_energy(LsCtl* c, const double* score, int is_spin_glass) {
    c->max_energy = is_spin_glass ? score[1] : score[0];
}

// Start of a local search over int8 weights with adjacency lists, in one work
// group: the state from the caller's 0/1 bytes (read from pinned host memory),
// the exact int32 field h = W s (k_field_adj's arithmetic), the exact score
// (k_score's), the reheat copy, the control record (max_energy = the start's
// score, then the first round's reset as k_ls_begin makes it) and cleared
// field-repair counters.  It replaces the copy, field, score, initialisation
// and reset launches of the general start.
__global__ __launch_bounds__(1024) void k_ls_start_i8(GView<int8_t> g, const uint8_t* theta, uint8_t* x, int32_t* h,
                                                     uint8_t* rx, int32_t* rh, LsCtl* c, FieldRepair* fix,
                                                     double scale, int64_t max_rounds, double w_up, int64_t w_up_i,
                                                     int is_spin_glass) {
    __shared__ long long part[16];
    const int64_t n = g.n;
    if ((n & 15) == 0 && (((uintptr_t)theta | (uintptr_t)x | (uintptr_t)rx) & 15) == 0) {
        // 16-byte pieces: one round of reads from host memory
        for (int64_t k = threadIdx.x; k < (n >> 4); k += blockDim.x) {
            const uint4 t = ((const uint4*)theta)[k];
            ((uint4*)x)[k] = t;
            ((uint4*)rx)[k] = t;
        }
    } else {
        for (int64_t v = threadIdx.x; v < n; v += blockDim.x) {
            const uint8_t xv = theta[v];
            x[v] = xv;
            rx[v] = xv;
        }
    }
    __syncthreads();
    long long q = 0;
    for (int64_t v = threadIdx.x; v < n; v += blockDim.x) {
        int32_t acc = 0;
        for (int32_t e = g.ap[v]; e < g.ap[v + 1]; ++e) acc += x[g.ai[e]] ? -(int32_t)g.aw[e] : (int32_t)g.aw[e];
        h[v] = acc;
        rh[v] = acc;
        q += x[v] ? -(long long)acc : (long long)acc;
    }
    q = wave_sum(q);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = q;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long qs = 0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) qs += part[i];
        (void)w_up;
        const double cut = (double)((2 * w_up_i - qs) / 4), energy = (double)(-qs / 2);
        LsCtl z = {};
        z.scale = scale;
        z.phase = -1;
        z.max_rounds = max_rounds;
        z.max_energy = is_spin_glass ? energy : cut;
        // the first round's reset (k_ls_begin)
        z.rounds = 1;
        z.skip_seed = 1;
        *c = z;
        ls_mark(c, 0);
        fix->checks = 0;
        fix->entries = 0;
        fix->max_bits = 0;
    }
}

void launch_ls_start_i8(const Graph& g, const uint8_t* theta_dev, uint8_t* x, void* h, uint8_t* rx, void* rh,
                        LsCtl* c, FieldRepair* fix, double scale, int64_t max_rounds, int is_spin_glass,
                        hipStream_t s) {
    hipLaunchKernelGGL(k_ls_start_i8, dim3(1), dim3(1024), 0, s, view<int8_t>(g), theta_dev, x, (int32_t*)h, rx,
                       (int32_t*)rh, c, fix, scale, max_rounds, g.w_up, g.w_up_i, is_spin_glass);
}

// End of an iteration of the int8 fast path, in one work group: the control
// record, the repair counters, the exact score of the state (k_score's
// arithmetic on the exact int32 field) and the state are written into a report
// in pinned host memory, then the next iteration's control reset (k_ls_begin)
// is applied.  The report is published by its sequence number: wave 0 writes
// the state and the report, then a system-scope fence, then `seq` (one more
// than the previous report's, counted in device memory), so a host that reads
// the expected number reads a whole report.
__device__ void ls_report_body(const uint8_t* x, const int32_t* h, int64_t n, int64_t w_up_i, LsCtl* c,
                               const FieldRepair* fix, LsReport* rep, uint8_t* x_out, unsigned long long* seq) {
    __shared__ long long part[16];
    long long q = 0;
    for (int64_t v = threadIdx.x; v < n; v += blockDim.x) {
        const uint8_t xv = x[v];
        q += xv ? -(long long)h[v] : (long long)h[v];
    }
    long long xs = 0;  // the state's terms of the check word (wave 0)
    if (threadIdx.x < 64) {
        // 16-byte pieces, then the ragged end byte by byte (x and x_out are 256-byte aligned)
        const int64_t nv = n >> 4;
        for (int64_t k = threadIdx.x; k < nv; k += 64) ((uint4*)x_out)[k] = ((const uint4*)x)[k];
        for (int64_t v = (nv << 4) + threadIdx.x; v < n; v += 64) x_out[v] = x[v];
        const int64_t n8 = n >> 3;
        for (int64_t j = threadIdx.x; j < n8; j += 64)
            xs += (long long)ls_state_term(((const unsigned long long*)x)[j], (unsigned long long)j);
        if (threadIdx.x == 0 && (n & 7)) {
            unsigned long long t = 0;
            for (int64_t v = n8 << 3; v < n; ++v) t |= (unsigned long long)x[v] << (8 * (v - (n8 << 3)));
            xs += (long long)ls_state_term(t, (unsigned long long)n8);
        }
    }
    q = wave_sum(q);
    xs = wave_sum(xs);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = q;
    __syncthreads();
    if (threadIdx.x < 64) {
        if (threadIdx.x == 0) {
            long long qs = 0;
            for (int i = 0; i < (int)(blockDim.x >> 6); ++i) qs += part[i];
            LsReport loc;
            loc.ctl = *c;
            loc.fix = *fix;
            loc.cut = (double)((2 * w_up_i - qs) / 4);
            loc.energy = (double)(-qs / 2);
            rep->cut = loc.cut;
            rep->energy = loc.energy;
            rep->ctl = loc.ctl;
            rep->fix = loc.fix;
            rep->check = ls_report_check(loc, (unsigned long long)xs, *seq + 1);
        }
        // A fence orders only its own thread's stores: every thread that
        // stored part of the state or the report (wave 0) fences them at
        // system scope, and the barrier below puts all of that before lane
        // 0's store of the number.  Moving the copy to other threads needs
        // them fenced here too.
        __threadfence_system();
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long k = *seq + 1;
        *seq = k;
        __hip_atomic_store(&rep->seq, k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        ls_begin_body(c);
    }
}

__global__ __launch_bounds__(1024) void k_ls_report_i8(const uint8_t* x, const int32_t* h, int64_t n, int64_t w_up_i,
                                                      LsCtl* c, const FieldRepair* fix, LsReport* rep,
                                                      uint8_t* x_out, unsigned long long* seq) {
    ls_report_body(x, h, n, w_up_i, c, fix, rep, x_out, seq);
}

void launch_ls_report_i8(const Graph& g, const uint8_t* x, const void* h, LsCtl* c, const FieldRepair* fix,
                         LsReport* rep_dev, uint8_t* x_out_dev, unsigned long long* seq, hipStream_t s) {
    hipLaunchKernelGGL(k_ls_report_i8, dim3(1), dim3(1024), 0, s, x, (const int32_t*)h, g.n, g.w_up_i, c, fix,
                       rep_dev, x_out_dev, seq);
}

void launch_ls_start_energy(LsCtl* c, const double* score, int is_spin_glass, hipStream_t s) {
    hipLaunchKernelGGL(k_ls_start_energy, dim3(1), dim3(1), 0, s, c, score, is_spin_glass);
}

void launch_ls_after_descent(LsCtl* c, const PhaseOut* o, int64_t cap, hipStream_t s) {
    hipLaunchKernelGGL(k_ls_after_descent, dim3(1), dim3(1), 0, s, c, o, cap);
}

template <typename WT>
static void ls_after_double_impl(const Graph& g, LsCtl* c, const PhaseOut* o, uint8_t* x, void* h, int64_t evals,
                                 hipStream_t s) {
    using FT = typename FieldOf<WT>::T;
    hipLaunchKernelGGL((k_ls_after_double<WT, FT>), dim3(1), dim3(1024), 0, s, view<WT>(g), c, o, x, (FT*)h, evals);
}
void launch_ls_after_double(const Graph& g, LsCtl* c, const PhaseOut* o, uint8_t* x, void* h, int64_t evals,
                            hipStream_t s) {
    PQI_DISPATCH(g, ls_after_double_impl, g, c, o, x, h, evals, s);
}

template <typename WT>
static void ls_descent_double_impl(const Graph& g, LsCtl* c, const PhaseOut* dout, int64_t cap, uint8_t* x, void* h,
                                   int64_t thread_count, PhaseOut* dout2, int64_t evals, const SeedRows* sr,
                                   hipStream_t s) {
    using FT = typename FieldOf<WT>::T;
    const SeedRows rows = sr ? *sr : SeedRows{};
    const unsigned blocks = 1u + (unsigned)((rows.count * rows.wpr + 1023) / 1024);
    hipLaunchKernelGGL((k_ls_descent_double<WT, FT>), dim3(blocks), dim3(1024), 0, s, view<WT>(g), c, dout, cap, x,
                       (FT*)h, thread_count, dout2, evals, rows);
}
void launch_ls_descent_double(const Graph& g, LsCtl* c, const PhaseOut* dout, int64_t cap, uint8_t* x, void* h,
                              int64_t thread_count, PhaseOut* dout2, int64_t evals, hipStream_t s,
                              const SeedRows* sr) {
    PQI_DISPATCH(g, ls_descent_double_impl, g, c, dout, cap, x, h, thread_count, dout2, evals, sr, s);
}

template <typename WT>
static void ls_seed_take_impl(const Graph& g, LsCtl* c, double* seld, int64_t count, int64_t* sel, int64_t bs,
                              uint8_t* x, void* h, const unsigned long long* qi, int64_t G, const int32_t* rank,
                              hipStream_t s) {
    using FT = typename FieldOf<WT>::T;
    hipLaunchKernelGGL((k_ls_seed_take<WT, FT>), dim3(1), dim3(1024), 0, s, view<WT>(g), c, seld, count, sel, bs, x,
                       (FT*)h, qi, G, rank);
}
void launch_ls_seed_take(const Graph& g, LsCtl* c, double* seld, int64_t count, int64_t* sel, int64_t bs,
                         uint8_t* x, void* h, hipStream_t s, const unsigned long long* qi, int64_t G,
                         const int32_t* rank) {
    PQI_DISPATCH(g, ls_seed_take_impl, g, c, seld, count, sel, bs, x, h, qi, G, rank, s);
}

template <typename WT>
static void ls_walk_reheat_impl(const Graph& g, LsCtl* c, const double* we, const int64_t* wev, int64_t walkers,
                                int64_t gray_evals, const uint8_t* wx, const void* wh, uint8_t* x, void* h,
                                uint8_t* rx, void* rh, PhaseOut* dout, int64_t thread_count, PhaseOut* dout2,
                                int64_t dbl_evals, const LsReportTo* rt, hipStream_t s) {
    using FT = typename FieldOf<WT>::T;
    LsReportTo to{};
    if (rt && std::is_same<FT, int32_t>::value) to = *rt;
    hipLaunchKernelGGL((k_ls_walk_reheat<WT, FT>), dim3(1), dim3(1024), 0, s, view<WT>(g), c, we, wev, walkers,
                       gray_evals, wx, (const FT*)wh, x, (FT*)h, rx, (FT*)rh, dout, thread_count, dout2, dbl_evals,
                       to);
}
void launch_ls_walk_reheat(const Graph& g, LsCtl* c, const double* we, const int64_t* wev, int64_t walkers,
                           int64_t gray_evals, const uint8_t* wx, const void* wh, uint8_t* x, void* h, uint8_t* rx,
                           void* rh, PhaseOut* dout, int64_t thread_count, PhaseOut* dout2, int64_t dbl_evals,
                           hipStream_t s, const LsReportTo* rt) {
    PQI_DISPATCH(g, ls_walk_reheat_impl, g, c, we, wev, walkers, gray_evals, wx, wh, x, h, rx, rh, dout, thread_count,
                 dout2, dbl_evals, rt, s);
}

template <typename WT>
static void ls_after_walk_impl(const Graph& g, LsCtl* c, const double* we, const int64_t* wev, int64_t walkers,
                               int64_t gray_evals, const uint8_t* wx, const void* wh, uint8_t* x, void* h,
                               uint8_t* rx, void* rh, hipStream_t s) {
    using FT = typename FieldOf<WT>::T;
    hipLaunchKernelGGL((k_ls_after_walk<WT, FT>), dim3(1), dim3(1024), 0, s, view<WT>(g), c, we, wev, walkers,
                       gray_evals, wx, (const FT*)wh, x, (FT*)h, rx, (FT*)rh);
}
void launch_ls_after_walk(const Graph& g, LsCtl* c, const double* we, const int64_t* wev, int64_t walkers,
                          int64_t gray_evals, const uint8_t* wx, const void* wh, uint8_t* x, void* h, uint8_t* rx,
                          void* rh, hipStream_t s) {
    PQI_DISPATCH(g, ls_after_walk_impl, g, c, we, wev, walkers, gray_evals, wx, wh, x, h, rx, rh, s);
}

template <typename WT>
static void ls_reheat_single_impl(const Graph& g, LsCtl* c, const PhaseOut* o, uint8_t* x, void* h, uint8_t* rx,
                                  void* rh, hipStream_t s) {
    using FT = typename FieldOf<WT>::T;
    hipLaunchKernelGGL((k_ls_reheat_single<WT, FT>), dim3(1), dim3(1024), 0, s, view<WT>(g), c, o, x, (FT*)h, rx,
                       (FT*)rh);
}
void launch_ls_reheat_single(const Graph& g, LsCtl* c, const PhaseOut* o, uint8_t* x, void* h, uint8_t* rx,
                             void* rh, hipStream_t s) {
    PQI_DISPATCH(g, ls_reheat_single_impl, g, c, o, x, h, rx, rh, s);
}

template <typename WT>
static void ls_reheat_double_impl(const Graph& g, LsCtl* c, const PhaseOut* o, uint8_t* x, void* h, uint8_t* rx,
                                  void* rh, int64_t evals, hipStream_t s) {
    using FT = typename FieldOf<WT>::T;
    hipLaunchKernelGGL((k_ls_reheat_double<WT, FT>), dim3(1), dim3(1024), 0, s, view<WT>(g), c, o, x, (FT*)h, rx,
                       (FT*)rh, evals);
}
void launch_ls_reheat_double(const Graph& g, LsCtl* c, const PhaseOut* o, uint8_t* x, void* h, uint8_t* rx,
                             void* rh, int64_t evals, hipStream_t s) {
    PQI_DISPATCH(g, ls_reheat_double_impl, g, c, o, x, h, rx, rh, evals, s);
}

// ===========================================================================
// Batched search: the same phase bodies over a list of instances, one work
// group (or a row of work groups) per listed instance.  Per-instance pointers
// come from BInstRaw descriptors in device memory.
// ===========================================================================

template <typename WT>
__device__ __forceinline__ GView<WT> gview_of(const BInstRaw& r) {
    GView<WT> v;
    v.W = (const WT*)r.W;
    v.ld = r.ld;
    v.n = r.n;
    v.ap = r.ap;
    v.ai = r.ai;
    v.aw = (const WT*)r.aw;
    v.use_adj = r.use_adj;
    v.kadj = r.kadj;
    v.kstride = r.kadj ? r.kstride : 0;
    return v;
}

// every listed instance sparse: the round descent in a kernel of its own
template <typename WT, typename FT>
__global__ __launch_bounds__(64) void kb_descent_round(const BInstRaw* insts, const int32_t* list, int64_t max_steps,
                                                      int which) {
    extern __shared__ __align__(16) unsigned char lds[];
    const BInstRaw& r = insts[list[blockIdx.x]];
    descent_round_body<WT, FT>(gview_of<WT>(r), which ? r.rx : r.x, (FT*)(which ? r.rh : r.h), max_steps, r.out, lds);
}

// mode 2: the round descent for sparse instances (lists mixing sparse and
// dense instances), 1 / 0 as for k_descent_wave
template <typename WT, typename FT>
__global__ __launch_bounds__(64) void kb_descent_lds(const BInstRaw* insts, const int32_t* list, int64_t max_steps,
                                                    int which, int mode) {
    extern __shared__ __align__(16) unsigned char lds[];
    const BInstRaw& r = insts[list[blockIdx.x]];
    if (mode == 2 && r.use_adj)
        descent_round_body<WT, FT>(gview_of<WT>(r), which ? r.rx : r.x, (FT*)(which ? r.rh : r.h), max_steps, r.out,
                                    lds);
    else
        descent_wave_body<WT, FT>(gview_of<WT>(r), which ? r.rx : r.x, (FT*)(which ? r.rh : r.h), max_steps, r.out,
                                  lds, mode != 0 && r.use_adj);
}

__global__ __launch_bounds__(64) void kb_descent_key(const BInstRaw* insts, const int32_t* list, int64_t max_steps,
                                                    int which, size_t lds_bytes) {
    extern __shared__ __align__(16) unsigned char lds[];
    const BInstRaw& r = insts[list[blockIdx.x]];
    descent_key_body(gview_of<int8_t>(r), which ? r.rx : r.x, (int32_t*)(which ? r.rh : r.h), max_steps, r.out, lds,
                     lds_bytes, r.key_shift);
}

size_t b_key_lds_limit() {
    static const size_t limit = lds_limit((const void*)kb_descent_key);
    return limit;
}

template <int R, int NB>
__global__ __launch_bounds__(64) void kb_descent_reg(const BInstRaw* insts, const int32_t* list, int64_t max_steps,
                                                    int which, size_t lds_bytes) {
    extern __shared__ __align__(16) unsigned char lds[];
    const BInstRaw& r = insts[list[blockIdx.x]];
    descent_reg_body<R, NB>(gview_of<int8_t>(r), which ? r.rx : r.x, (int32_t*)(which ? r.rh : r.h), max_steps,
                            r.out, lds, lds_bytes, r.key_shift);
}

template <int NW, int NP, int NB, int H>
__global__ __launch_bounds__(64) void kb_descent_plane(const BInstRaw* insts, const int32_t* list, int64_t max_steps,
                                                      int which, size_t lds_bytes, int S) {
    extern __shared__ __align__(16) unsigned char lds[];
    const BInstRaw& r = insts[list[blockIdx.x]];
    descent_plane_body<NW, NP, NB, H>(gview_of<int8_t>(r), which ? r.rx : r.x, (int32_t*)(which ? r.rh : r.h),
                                      max_steps, r.out, lds, lds_bytes, S);
}

template <int NW, int NP, int NB, int H>
static void launch_b_plane(const BInstRaw* d, const int32_t* list, int64_t L, int64_t max_steps, int which,
                           size_t table, int S, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    hipExtLaunchKernelGGL((kb_descent_plane<NW, NP, NB, H>), dim3((unsigned)L), dim3(64), table, s, e0, e1, 0, d,
                          list, max_steps, which, table, S);
}

template <int R, int NB>
static void launch_b_reg(const BInstRaw* d, const int32_t* list, int64_t L, int64_t max_steps, int which,
                         size_t table, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    static const size_t limit = lds_limit((const void*)kb_descent_reg<R, NB>);
    (void)limit;  // the plan admits only tables within it
    hipExtLaunchKernelGGL((kb_descent_reg<R, NB>), dim3((unsigned)L), dim3(64), table, s, e0, e1, 0, d, list,
                          max_steps, which, table);
}

template <typename WT, typename FT>
__global__ __launch_bounds__(64) void kb_descent_lds2(const BInstRaw* insts, const int32_t* list, int64_t max_steps,
                                                     int which) {
    extern __shared__ __align__(16) unsigned char lds[];
    const BInstRaw& r = insts[list[blockIdx.x]];
    descent_wave2_body<WT, FT>(gview_of<WT>(r), which ? r.rx : r.x, (FT*)(which ? r.rh : r.h), max_steps, r.out, lds);
}

template <typename WT, typename FT>
__global__ __launch_bounds__(1024) void kb_descent_glb(const BInstRaw* insts, const int32_t* list,
                                                      int64_t max_steps, int which) {
    const BInstRaw& r = insts[list[blockIdx.x]];
    descent_body<WT, FT>(gview_of<WT>(r), which ? r.rx : r.x, (FT*)(which ? r.rh : r.h), max_steps, r.out);
}

template <typename WT, typename FT>
__global__ __launch_bounds__(1024) void kb_double(const BInstRaw* insts, const int32_t* list, int which) {
    const BInstRaw& r = insts[list[blockIdx.x]];
    double_ref_body<WT, FT>(gview_of<WT>(r), which ? r.rx : r.x, (const FT*)(which ? r.rh : r.h), r.G, r.out);
}

template <typename WT, typename FT>
__global__ __launch_bounds__(1024) void kb_flip(const BInstRaw* insts, const int32_t* list, const int64_t* moves,
                                               int which) {
    const BInstRaw& r = insts[list[blockIdx.x]];
    flip_body<WT, FT>(gview_of<WT>(r), which ? r.rx : r.x, (FT*)(which ? r.rh : r.h), moves[2 * blockIdx.x],
                      moves[2 * blockIdx.x + 1]);
}

template <typename WT, typename FT>
__global__ __launch_bounds__(256) void kb_seed_lowrank(const BInstRaw* insts, const int32_t* list) {
    const BInstRaw& r = insts[list[blockIdx.y]];
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b < r.seed_count) r.delta[b] = seed_lowrank_one<WT, FT>(gview_of<WT>(r), r.x, (const FT*)r.h, r.bs, b);
    if (b < r.seed_count) r.rank[b] = 0;
}

__global__ __launch_bounds__(256) void kb_rank(const BInstRaw* insts, const int32_t* list, int64_t splits) {
    const BInstRaw& r = insts[list[blockIdx.z]];
    if ((int64_t)blockIdx.x * blockDim.x >= r.seed_count) return;  // whole work group idle
    const int64_t span = (r.seed_count + splits - 1) / splits;
    if ((int64_t)blockIdx.y * span >= r.seed_count) return;
    rank_body(r.delta, r.seed_count, span, r.rank, blockIdx.x, blockIdx.y);
}

__global__ void kb_rank_scatter(const BInstRaw* insts, const int32_t* list) {
    const BInstRaw& r = insts[list[blockIdx.y]];
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a < r.seed_count && r.rank[a] < r.G) r.sel[r.rank[a]] = a;
}

// Stable top G of the seed gains, one 1024-thread work group per instance:
// G rounds of a block first-argmax (largest gain, then lowest index), each
// taking the winner out.  The winners come in rank order, the order of the
// rank counting below (descending gain, ties by ascending index), so sel and
// seld are the same.  Thread t owns seeds t, t + 1024, ... (at most 64 of
// them: seed_count <= 65536), remembers the ones taken in a bit mask, and
// rescans its own seeds only when it supplied the winner.
constexpr int64_t kTopkMaxSeeds = 64 * 1024;

__global__ __launch_bounds__(1024) void kb_topk(const BInstRaw* insts, const int32_t* list) {
    __shared__ double sv[16];
    __shared__ int32_t sk[16];
    const BInstRaw& r = insts[list[blockIdx.x]];
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    const int count = (int)r.seed_count;
    uint64_t taken = 0;
    double bv;
    int bk;
    auto rescan = [&]() {
        bv = -INFINITY;
        bk = INT_MAX;
        for (int j = 0; t + j * 1024 < count; ++j) {
            const int a = t + j * 1024;
            const double v = r.delta[a];
            if (!((taken >> j) & 1u) && (v > bv || bk == INT_MAX)) {
                bv = v;
                bk = a;
            }
        }
    };
    rescan();
    for (int64_t i = 0; i < r.G; ++i) {
        double v = bv;
        int k = bk;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_xor(v, off, 64);
            const int ok = __shfl_xor(k, off, 64);
            if (ov > v || (ov == v && ok < k)) {
                v = ov;
                k = ok;
            }
        }
        if (lane == 0) {
            sv[wid] = v;
            sk[wid] = k;
        }
        __syncthreads();
        v = sv[0];
        k = sk[0];
        for (int w = 1; w < 16; ++w)
            if (sv[w] > v || (sv[w] == v && sk[w] < k)) {
                v = sv[w];
                k = sk[w];
            }
        __syncthreads();
        if (t == 0) {
            r.sel[i] = k;
            r.seld[i] = v;
        }
        if (k != INT_MAX && (k & 1023) == t) {
            taken |= 1ull << (k >> 10);
            rescan();
        }
    }
}

__global__ void kb_gather(const BInstRaw* insts, const int32_t* list) {
    const BInstRaw& r = insts[list[blockIdx.y]];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < r.G) r.seld[i] = r.delta[r.sel[i]];
}

template <typename WT, typename FT>
__global__ __launch_bounds__(64) void kb_walk(const BInstRaw* insts, const int32_t* list, int64_t T_override,
                                             int64_t G_override, int use_lds) {
    extern __shared__ __align__(16) unsigned char lds[];
    const BInstRaw& r = insts[list[blockIdx.y]];
    const int64_t G = G_override > 0 ? G_override : r.G;
    if ((int64_t)blockIdx.x >= G) return;
    const int64_t T = T_override >= 0 ? T_override : r.T;
    gray_walk_body<WT, FT>(gview_of<WT>(r), r.x, (const FT*)r.h, r.sel, r.seld, r.bs, T, r.wx, (FT*)r.wh, r.we,
                           r.wev, use_lds, lds, blockIdx.x);
}

template <int NB>
__global__ __launch_bounds__(64) void kb_walk_reg(const BInstRaw* insts, const int32_t* list, int64_t T_override,
                                                 int64_t G_override, size_t lds_bytes) {
    extern __shared__ __align__(16) unsigned char lds[];
    const BInstRaw& r = insts[list[blockIdx.y]];
    const int64_t G = G_override > 0 ? G_override : r.G;
    if ((int64_t)blockIdx.x >= G) return;
    const int64_t T = T_override >= 0 ? T_override : r.T;
    gray_walk_reg_body<NB>(gview_of<int8_t>(r), r.x, (const int32_t*)r.h, r.sel, r.seld, r.bs, T, r.wx,
                           (int32_t*)r.wh, r.we, r.wev, lds, lds_bytes, blockIdx.x);
}

template <int NB>
static void launch_b_walk_reg(const BInstRaw* d, const int32_t* list, int64_t L, int64_t gx, int64_t T_override,
                              int64_t G_override, size_t table, hipStream_t s) {
    static const size_t limit = lds_limit((const void*)kb_walk_reg<NB>);
    const size_t lds = table <= limit ? table : 0;
    hipLaunchKernelGGL(kb_walk_reg<NB>, dim3((unsigned)gx, (unsigned)L), dim3(64), lds, s, d, list, T_override,
                       G_override, lds);
}

// copy a state: src >= 0 -> walker src, src == -1 -> reheat state; dst 0 -> best, 1 -> reheat
template <typename FT>
__global__ __launch_bounds__(256) void kb_copy(const BInstRaw* insts, const int32_t* list, const int32_t* src,
                                              int dst) {
    const BInstRaw& r = insts[list[blockIdx.y]];
    const int sidx = src[blockIdx.y];
    const uint8_t* sx = sidx >= 0 ? r.wx + (int64_t)sidx * r.n : r.rx;
    const FT* sh = sidx >= 0 ? (const FT*)r.wh + (int64_t)sidx * r.n : (const FT*)r.rh;
    uint8_t* dx = dst ? r.rx : r.x;
    FT* dh = (FT*)(dst ? r.rh : r.h);
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < r.n; v += (int64_t)gridDim.x * blockDim.x) {
        dx[v] = sx[v];
        dh[v] = sh[v];
    }
}

template <typename FT>
__global__ __launch_bounds__(1024) void kb_quad(const BInstRaw* insts, const int32_t* list, double* q) {
    const BInstRaw& r = insts[list[blockIdx.x]];
    quad_body<FT>(r.x, (const FT*)r.h, r.n, q + blockIdx.x);
}

// Same summation order as k_field_dense / k_field_adj, so every instance's
// field is bit-identical to the single-instance path.
template <typename WT, typename FT>
__global__ __launch_bounds__(256) void kb_field(const BInstRaw* insts, const int32_t* list) {
    const BInstRaw& r = insts[list[blockIdx.y]];
    const GView<WT> g = gview_of<WT>(r);
    FT* h = (FT*)r.h;
    if (g.use_adj) {
        const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (row >= r.n) return;
        FT acc = 0;
        for (int32_t e = g.ap[row]; e < g.ap[row + 1]; ++e) acc += r.x[g.ai[e]] ? -(FT)g.aw[e] : (FT)g.aw[e];
        h[row] = acc;
    } else {
        const int lane = threadIdx.x & 63;
        for (int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); row < r.n; row += (int64_t)gridDim.x * 4) {
            const WT* Wr = g.W + row * g.ld;
            FT acc = 0;
            for (int64_t u = lane; u < r.n; u += 64) acc += r.x[u] ? -(FT)Wr[u] : (FT)Wr[u];
            acc = wave_sum(acc);
            if (lane == 0) h[row] = acc;
        }
    }
}

// Sparse rows take a thread each (max_n / 256 work groups per instance);
// dense rows a wave each, in a loop over the grid, so a grid sized for
// sparse rows also covers dense instances.
template <typename WT>
static void b_field_impl(const BInstRaw* d, const int32_t* list, int64_t L, int64_t max_n, int all_sparse,
                         hipStream_t s) {
    using FT = typename FieldOf<WT>::T;
    const int64_t gx = all_sparse ? (max_n + 255) / 256 : (max_n + 3) / 4;
    hipLaunchKernelGGL((kb_field<WT, FT>), dim3((unsigned)gx, (unsigned)L), dim3(256), 0, s, d, list);
}

template <typename WT>
static void b_descent_impl(const BInstRaw* d, const int32_t* list, int64_t L, int64_t max_n, int64_t max_steps,
                           int which, int all_sparse, size_t key_lds, int reg_r, int64_t reg_stride, size_t reg_lds,
                           int plane_s, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    using FT = typename FieldOf<WT>::T;
    if (sizeof(WT) == 1 && reg_r > 0 && plane_s >= 0) {
        const int S = plane_s >> 1, half = plane_s & 1;
        PQI_PLANE_DISPATCH(max_n <= 2048 ? 1 : 2, plane_class(S), plane_h(half, S), launch_b_plane, d, list, L,
                           max_steps, which, reg_lds, S, s, e0, e1);
        return;
    }
    if (sizeof(WT) == 1 && reg_r > 0) {
        PQI_REG_DISPATCH(reg_r, reg_nb(reg_stride), launch_b_reg, d, list, L, max_steps, which, reg_lds, s, e0, e1);
        return;
    }
    if (sizeof(WT) == 1 && key_lds > 0 && descent_variant() != 3) {
        hipExtLaunchKernelGGL(kb_descent_key, dim3((unsigned)L), dim3(64), key_lds, s, e0, e1, 0, d, list, max_steps,
                              which, key_lds);
        return;
    }
    static const size_t limit = std::min(lds_limit((const void*)kb_descent_lds<WT, FT>),
                                         lds_limit((const void*)kb_descent_round<WT, FT>));
    static const size_t limit2 = lds_limit((const void*)kb_descent_lds2<WT, FT>);
    const bool stage_ap = descent_variant() != 7 && wave_lds<FT>(max_n, true) <= limit;
    const bool rounds = round_fits<WT, FT>(max_n, limit);
    const size_t lds = rounds ? round_lds_layout<WT, FT>(max_n).total : wave_lds<FT>(max_n, stage_ap);
    if (all_sparse && descent_variant() == 2 && descent2_lds<FT>(max_n) <= limit2)
        hipExtLaunchKernelGGL((kb_descent_lds2<WT, FT>), dim3((unsigned)L), dim3(64), descent2_lds<FT>(max_n), s, e0,
                              e1, 0, d, list, max_steps, which);
    else if (rounds && all_sparse)
        hipExtLaunchKernelGGL((kb_descent_round<WT, FT>), dim3((unsigned)L), dim3(64), lds, s, e0, e1, 0, d, list,
                              max_steps, which);
    else if (lds <= limit)
        hipExtLaunchKernelGGL((kb_descent_lds<WT, FT>), dim3((unsigned)L), dim3(64), lds, s, e0, e1, 0, d, list,
                              max_steps, which, rounds ? 2 : stage_ap ? 1 : 0);
    else
        hipExtLaunchKernelGGL((kb_descent_glb<WT, FT>), dim3((unsigned)L), dim3(1024), 0, s, e0, e1, 0, d, list,
                              max_steps, which);
}

template <typename WT>
static void b_double_impl(const BInstRaw* d, const int32_t* list, int64_t L, int which, hipStream_t s) {
    using FT = typename FieldOf<WT>::T;
    hipLaunchKernelGGL((kb_double<WT, FT>), dim3((unsigned)L), dim3(1024), 0, s, d, list, which);
}

template <typename WT>
static void b_flip_impl(const BInstRaw* d, const int32_t* list, int64_t L, const int64_t* moves, int which,
                        hipStream_t s) {
    using FT = typename FieldOf<WT>::T;
    hipLaunchKernelGGL((kb_flip<WT, FT>), dim3((unsigned)L), dim3(1024), 0, s, d, list, moves, which);
}

template <typename WT>
static void b_seeds_impl(const BInstRaw* d, const int32_t* list, int64_t L, int64_t max_seeds, int64_t max_G,
                         hipStream_t s) {
    using FT = typename FieldOf<WT>::T;
    const unsigned gx = (unsigned)((max_seeds + 255) / 256);
    hipLaunchKernelGGL((kb_seed_lowrank<WT, FT>), dim3(gx, (unsigned)L), dim3(256), 0, s, d, list);
    if (max_seeds <= kTopkMaxSeeds) {
        hipLaunchKernelGGL(kb_topk, dim3((unsigned)L), dim3(1024), 0, s, d, list);
        return;
    }
    int64_t splits = std::min<int64_t>(64, (max_seeds + 511) / 512);
    hipLaunchKernelGGL(kb_rank, dim3(gx, (unsigned)splits, (unsigned)L), dim3(256), 0, s, d, list, splits);
    hipLaunchKernelGGL(kb_rank_scatter, dim3(gx, (unsigned)L), dim3(256), 0, s, d, list);
    hipLaunchKernelGGL(kb_gather, dim3((unsigned)((max_G + 255) / 256), (unsigned)L), dim3(256), 0, s, d, list);
}

template <typename WT>
static void b_walk_impl(const BInstRaw* d, const int32_t* list, int64_t L, int64_t max_n, int64_t max_G,
                        int64_t T_override, int64_t G_override, int reg_r, int64_t reg_stride, size_t reg_lds,
                        hipStream_t s) {
    using FT = typename FieldOf<WT>::T;
    const int64_t gx = G_override > 0 ? G_override : max_G;
    if (sizeof(WT) == 1 && reg_r > 0 && max_n <= 2048) {
        switch (reg_nb(reg_stride)) {
            case 4: launch_b_walk_reg<4>(d, list, L, gx, T_override, G_override, reg_lds, s); break;
            case 8: launch_b_walk_reg<8>(d, list, L, gx, T_override, G_override, reg_lds, s); break;
            default: launch_b_walk_reg<16>(d, list, L, gx, T_override, G_override, reg_lds, s); break;
        }
        return;
    }
    const size_t lds = (((size_t)max_n * sizeof(FT) + 15) & ~(size_t)15) + (size_t)max_n;
    static const size_t limit = lds_limit((const void*)kb_walk<WT, FT>);
    const int use_lds = lds <= limit ? 1 : 0;
    hipLaunchKernelGGL((kb_walk<WT, FT>), dim3((unsigned)gx, (unsigned)L), dim3(64), use_lds ? lds : 0, s, d, list,
                       T_override, G_override, use_lds);
}

template <typename WT>
static void b_copy_impl(const BInstRaw* d, const int32_t* list, int64_t L, int64_t max_n, const int32_t* src, int dst,
                        hipStream_t s) {
    using FT = typename FieldOf<WT>::T;
    hipLaunchKernelGGL((kb_copy<FT>), dim3((unsigned)std::min<int64_t>(64, (max_n + 255) / 256), (unsigned)L),
                       dim3(256), 0, s, d, list, src, dst);
}

template <typename WT>
static void b_quad_impl(const BInstRaw* d, const int32_t* list, int64_t L, double* q, hipStream_t s) {
    using FT = typename FieldOf<WT>::T;
    hipLaunchKernelGGL((kb_quad<FT>), dim3((unsigned)L), dim3(1024), 0, s, d, list, q);
}

#define PQI_KIND_DISPATCH(kind, FN, ...)                            \
    do {                                                            \
        switch (kind) {                                             \
            case PQI_WEIGHTS_I8: FN<int8_t>(__VA_ARGS__); break;    \
            case PQI_WEIGHTS_F32: FN<float>(__VA_ARGS__); break;    \
            default: FN<double>(__VA_ARGS__); break;                \
        }                                                           \
    } while (0)

void launch_b_field(int kind, const BInstRaw* d, const int32_t* list, int64_t L, int64_t max_n, int all_sparse,
                    hipStream_t s) {
    PQI_KIND_DISPATCH(kind, b_field_impl, d, list, L, max_n, all_sparse, s);
}
void launch_b_descent(int kind, const BInstRaw* d, const int32_t* list, int64_t L, int64_t max_n, int64_t max_steps,
                      int which, int all_sparse, size_t key_lds, int reg_r, int64_t reg_stride, size_t reg_lds,
                      int plane_s, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    PQI_KIND_DISPATCH(kind, b_descent_impl, d, list, L, max_n, max_steps, which, all_sparse, key_lds, reg_r,
                      reg_stride, reg_lds, plane_s, s, e0, e1);
}
void launch_b_double(int kind, const BInstRaw* d, const int32_t* list, int64_t L, int which, hipStream_t s) {
    PQI_KIND_DISPATCH(kind, b_double_impl, d, list, L, which, s);
}
void launch_b_flip(int kind, const BInstRaw* d, const int32_t* list, int64_t L, const int64_t* moves, int which,
                   hipStream_t s) {
    PQI_KIND_DISPATCH(kind, b_flip_impl, d, list, L, moves, which, s);
}
void launch_b_seeds(int kind, const BInstRaw* d, const int32_t* list, int64_t L, int64_t max_seeds, int64_t max_G,
                    hipStream_t s) {
    PQI_KIND_DISPATCH(kind, b_seeds_impl, d, list, L, max_seeds, max_G, s);
}
void launch_b_walk(int kind, const BInstRaw* d, const int32_t* list, int64_t L, int64_t max_n, int64_t max_G,
                   int64_t T_override, int64_t G_override, int reg_r, int64_t reg_stride, size_t reg_lds,
                   hipStream_t s) {
    PQI_KIND_DISPATCH(kind, b_walk_impl, d, list, L, max_n, max_G, T_override, G_override, reg_r, reg_stride, reg_lds,
                      s);
}
void launch_b_copy(int kind, const BInstRaw* d, const int32_t* list, int64_t L, int64_t max_n, const int32_t* src,
                   int dst, hipStream_t s) {
    PQI_KIND_DISPATCH(kind, b_copy_impl, d, list, L, max_n, src, dst, s);
}
void launch_b_quad(int kind, const BInstRaw* d, const int32_t* list, int64_t L, double* q, hipStream_t s) {
    PQI_KIND_DISPATCH(kind, b_quad_impl, d, list, L, q, s);
}

// ---------------------------------------------------------------------------
// Learned proposals
// ---------------------------------------------------------------------------

template <typename WT, typename FT>
__global__ __launch_bounds__(256) void kb_features(const BInstRaw* insts, const int32_t* list) {
    const BInstRaw& r = insts[list[blockIdx.y]];
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= r.n) return;
    const FT* h = (const FT*)r.h;
    const int sv = r.x[v] ? -1 : 1;
    const double d = (double)sv * (double)h[v];
    double str = 0.0;
    const WT* aw = (const WT*)r.aw;
    for (int32_t e = r.ap[v]; e < r.ap[v + 1]; ++e) str += fabs((double)aw[e]);
    float* f = r.feat + v * 4;
    f[0] = str > 0.0 ? (float)(d / str) : 0.f;
    f[1] = (float)(r.ap[v + 1] - r.ap[v]) * (1.f / 16.f);
    f[2] = (float)sv;
    f[3] = (float)(r.age[v] < 64 ? r.age[v] : 64) * (1.f / 64.f);
}

// Stable counting sort into 256 score buckets, highest scores first; within a
// bucket nodes keep index order.  One work group per instance.  Up to 8192
// nodes, each thread's 32 scores are read once, all in flight, and kept in
// registers for the range, the counts and the scatter.
__device__ void order_in_registers(const BInstRaw& r, int t) {
    constexpr int J = 32, KC = 4;
    __shared__ float lo_w[4], hi_w[4];
    __shared__ int32_t cnt[256];
    __shared__ int32_t wcnt[KC * 4][256];
    const int n = (int)r.n;
    const int lane = t & 63, wid = t >> 6;
    float sc[J];
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int v = t + 256 * j;
        sc[j] = r.score[v < n ? v : 0];
    }
    float lo = INFINITY, hi = -INFINITY;
#pragma unroll
    for (int j = 0; j < J; ++j)
        if (t + 256 * j < n) {
            lo = fminf(lo, sc[j]);
            hi = fmaxf(hi, sc[j]);
        }
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) {
        lo = fminf(lo, __shfl_xor(lo, m));
        hi = fmaxf(hi, __shfl_xor(hi, m));
    }
    if (lane == 0) {
        lo_w[wid] = lo;
        hi_w[wid] = hi;
    }
    cnt[t] = 0;
#pragma unroll
    for (int w = 0; w < KC * 4; ++w) wcnt[w][t] = 0;
    __syncthreads();
    lo = fminf(fminf(lo_w[0], lo_w[1]), fminf(lo_w[2], lo_w[3]));
    hi = fmaxf(fmaxf(hi_w[0], hi_w[1]), fmaxf(hi_w[2], hi_w[3]));
    const float span = hi > lo ? hi - lo : 1.f;
    auto bucket = [&](float x) {
        int b = (int)((hi - x) / span * 255.f);
        return b < 0 ? 0 : (b > 255 ? 255 : b);
    };
    int bk[J];
#pragma unroll
    for (int j = 0; j < J; ++j) {
        bk[j] = bucket(sc[j]);
        if (t + 256 * j < n) atomicAdd(&cnt[bk[j]], 1);
    }
    __syncthreads();
    if (t == 0) {
        int32_t run = 0;
        for (int b = 0; b < 256; ++b) {
            const int32_t c = cnt[b];
            cnt[b] = run;
            run += c;
        }
    }
    __syncthreads();
    // KC chunks of 256 nodes per round, as in the general path below
#pragma unroll
    for (int c = 0; c < J; c += KC) {
        if (256 * c >= n) break;  // uniform
        int rank[KC];
        bool ok[KC];
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            ok[k] = t + 256 * (c + k) < n;
            uint64_t same = __ballot(ok[k]);
#pragma unroll
            for (int bit = 0; bit < 8; ++bit) {
                const uint64_t m = __ballot(ok[k] && ((bk[c + k] >> bit) & 1));
                same &= ((bk[c + k] >> bit) & 1) ? m : ~m;
            }
            rank[k] = __popcll(same & ((1ull << lane) - 1ull));
            if (ok[k] && rank[k] == 0) wcnt[4 * k + wid][bk[c + k]] = __popcll(same);
        }
        __syncthreads();
        {
            int32_t run = cnt[t];
#pragma unroll
            for (int w = 0; w < KC * 4; ++w) {
                const int32_t x = wcnt[w][t];
                wcnt[w][t] = run;
                run += x;
            }
            cnt[t] = run;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < KC; ++k)
            if (ok[k]) r.perm[wcnt[4 * k + wid][bk[c + k]] + rank[k]] = t + 256 * (c + k);
        __syncthreads();
#pragma unroll
        for (int w = 0; w < KC * 4; ++w) wcnt[w][t] = 0;
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void kb_order(const BInstRaw* insts, const int32_t* list) {
    __shared__ float lo_s[256], hi_s[256];
    __shared__ int32_t cnt[256];
    const BInstRaw& r = insts[list[blockIdx.x]];
    const int t = threadIdx.x;
    if (r.n <= 8192) {
        order_in_registers(r, t);
        return;
    }
    float lo = INFINITY, hi = -INFINITY;
    for (int64_t v = t; v < r.n; v += 256) {
        const float sc = r.score[v];
        lo = fminf(lo, sc);
        hi = fmaxf(hi, sc);
    }
    lo_s[t] = lo;
    hi_s[t] = hi;
    cnt[t] = 0;
    __syncthreads();
    for (int m = 128; m > 0; m >>= 1) {
        if (t < m) {
            lo_s[t] = fminf(lo_s[t], lo_s[t + m]);
            hi_s[t] = fmaxf(hi_s[t], hi_s[t + m]);
        }
        __syncthreads();
    }
    lo = lo_s[0];
    hi = hi_s[0];
    const float span = hi > lo ? hi - lo : 1.f;
    auto bucket = [&](float sc) {
        int b = (int)((hi - sc) / span * 255.f);
        return b < 0 ? 0 : (b > 255 ? 255 : b);
    };
    for (int64_t v = t; v < r.n; v += 256) atomicAdd(&cnt[bucket(r.score[v])], 1);
    __syncthreads();
    // bucket starts (exclusive prefix of the counts)
    constexpr int KC = 4;  // 256-node chunks per round
    __shared__ int32_t wcnt[KC * 4][256];
    if (t == 0) {
        int32_t run = 0;
        for (int b = 0; b < 256; ++b) {
            const int32_t c = cnt[b];
            cnt[b] = run;
            run += c;
        }
    }
#pragma unroll
    for (int w = 0; w < KC * 4; ++w) wcnt[w][t] = 0;
    __syncthreads();
    // stable scatter, KC x 256 nodes in index order per round: in each chunk
    // the lanes of a wave with the same bucket are found by 8 ballots (one per
    // bucket bit) and ranked by lane; each bucket's offsets advance over the
    // (chunk, wave) groups in node order, so a round takes three barriers
    const int lane = t & 63, wid = t >> 6;
    for (int64_t c0 = 0; c0 < r.n; c0 += KC * 256) {
        int b[KC], rank[KC];
        bool ok[KC];
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            const int64_t v = c0 + 256 * k + t;
            ok[k] = v < r.n;
            b[k] = ok[k] ? bucket(r.score[v]) : 0;
        }
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            uint64_t same = __ballot(ok[k]);
#pragma unroll
            for (int bit = 0; bit < 8; ++bit) {
                const uint64_t m = __ballot(ok[k] && ((b[k] >> bit) & 1));
                same &= ((b[k] >> bit) & 1) ? m : ~m;
            }
            rank[k] = __popcll(same & ((1ull << lane) - 1ull));
            if (ok[k] && rank[k] == 0) wcnt[4 * k + wid][b[k]] = __popcll(same);
        }
        __syncthreads();
        {
            int32_t run = cnt[t];  // thread t: bucket t
#pragma unroll
            for (int w = 0; w < KC * 4; ++w) {
                const int32_t c = wcnt[w][t];
                wcnt[w][t] = run;
                run += c;
            }
            cnt[t] = run;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < KC; ++k)
            if (ok[k]) r.perm[wcnt[4 * k + wid][b[k]] + rank[k]] = (int32_t)(c0 + 256 * k + t);
        __syncthreads();
#pragma unroll
        for (int w = 0; w < KC * 4; ++w) wcnt[w][t] = 0;
        __syncthreads();
    }
}

// LDS of the staged learned pass: gains, spins, ages, row pointers and the
// visiting list
struct PassLds {
    size_t d, x, age, ap, list, total;
};

template <typename WT, typename FT>
__host__ __device__ inline PassLds pass_lds_layout(int64_t n) {
    auto al = [](size_t v) { return (v + 15) & ~(size_t)15; };
    PassLds L;
    size_t o = 0;
    L.d = o;
    o += al((size_t)(n + 1) * sizeof(FT));  // + a slot that idle lanes read and write back
    L.x = o;
    o += al((size_t)(n + 1));
    L.age = o;
    o += al((size_t)n);
    L.ap = o;
    o += al((size_t)(n + 1) * 4);
    L.list = o;
    o += al((size_t)n * 4);
    L.total = o;
    return L;
}

// The same pass with everything a flip touches in LDS or registers: the
// nodes whose score passes min_score are listed first (perm order), then
// each chunk of 64 listed nodes has its rows fetched into registers in one
// batch of loads before its flips, so a flip waits for no global access (rows
// longer than RD_RS come from global memory).  Ages stay in LDS until the end.
template <typename WT, typename FT>
__device__ void learned_pass_staged(const BInstRaw& r, const GView<WT>& g, float min_score, unsigned char* lds) {
    constexpr int RS = RD_RS;
    constexpr int WW = sizeof(WT) == 8 ? 2 : 1;  // 32-bit words per weight
    const int lane = threadIdx.x;
    const int n = (int)r.n;
    const PassLds L = pass_lds_layout<WT, FT>(n);
    FT* D = (FT*)(lds + L.d);
    uint8_t* x = lds + L.x;
    uint8_t* age = lds + L.age;
    int32_t* AP = (int32_t*)(lds + L.ap);
    int32_t* LIST = (int32_t*)(lds + L.list);
    FT* hg = (FT*)r.h;
    stage_gains<FT>(r.x, hg, n, x, D, nullptr);
    stage_words(g.ap, n + 1, AP);
    for (int v0 = lane; v0 < n; v0 += 64 * 16) {
        uint8_t a[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) a[u] = v0 + 64 * u < n ? r.age[v0 + 64 * u] : 0;
#pragma unroll
        for (int u = 0; u < 16; ++u)
            if (v0 + 64 * u < n) age[v0 + 64 * u] = a[u];
    }
    if (lane == 0) {
        D[n] = (FT)0;  // the idle lanes' slot
        x[n] = 0;
    }
    // the visiting list: perm order, score above min_score; 16 chunks of
    // perm and score loads in flight per lane
    int nl = 0;
    for (int c0 = 0; c0 < n; c0 += 64 * 16) {
        int pv[16];
        bool ok[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int i = c0 + 64 * u + lane;
            pv[u] = i < n ? r.perm[i] : -1;
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) ok[u] = pv[u] >= 0 && r.score[pv[u]] > min_score;
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const uint64_t m = __ballot(ok[u]);
            if (ok[u])
                LIST[nl + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))] = pv[u];
            nl += __popcll(m);
        }
    }
    __syncthreads();
    int64_t flips = 0;
    double gsum = 0.0, gmin = INFINITY;
    // the chunk's rows in registers: row i (the chunk's i-th node) has entry
    // j in register i >> 2 of lane 16 (i & 3) + j (node, or -1 past the row)
    int32_t ri[16];
    uint32_t rw0[16], rw1[WW == 2 ? 16 : 1];
    for (int c0 = 0; c0 < nl; c0 += 64) {
        {
            // every read of a stage in flight at once: no bounds tests
            // (indices clamped, results masked)
            int32_t cs[16], e0[16], e1[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int i = c0 + 4 * u + (lane >> 4);
                cs[u] = LIST[i < nl ? i : 0];
                cs[u] = i < nl ? cs[u] : -1;
            }
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int c = cs[u] >= 0 ? cs[u] : 0;
                e0[u] = AP[c];
                e1[u] = AP[c + 1];
            }
            int32_t vi[16];
            WT vw[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const bool ok = cs[u] >= 0 && (lane & 15) < e1[u] - e0[u];
                const int e = ok ? e0[u] + (lane & 15) : 0;
                vi[u] = g.ai[e];
                vw[u] = g.aw[e];
                e1[u] = ok;
            }
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                uint32_t w0, w1 = 0;
                weight_words(vw[u], w0, w1);
                ri[u] = e1[u] ? vi[u] : -1;
                rw0[u] = w0;
                if constexpr (WW == 2) rw1[u] = w1;
            }
        }
        // this lane's listed node: its spin, gain and row extent
        const bool valid = c0 + lane < nl;
        const int v = LIST[valid ? c0 + lane : 0];
        uint32_t xl = x[v];
        const int e0l = AP[v], degl = AP[v + 1] - e0l;
        FT d = valid ? D[v] : (FT)0;
        uint64_t mask = __ballot(valid && d > 0);
        while (mask) {
            const int bl = __builtin_ctzll(mask);
            const int k = __builtin_amdgcn_readlane(v, bl);
            const FT dd = readlane_t(d, bl);
            const int xk = (int)(__builtin_amdgcn_readlane(xl, bl) != 0);
            const int deg = __builtin_amdgcn_readlane(degl, bl);
            gsum += (double)dd;
            gmin = fmin(gmin, (double)dd);
            ++flips;
            if (deg <= RS) {
                // the row from registers: one entry per lane of its group; the
                // other lanes read and write back the idle slot n
                const int q = bl >> 2;
                int32_t ia = ri[q];
                uint32_t a0 = rw0[q], a1 = 0;
                if constexpr (WW == 2) a1 = rw1[q];
                __asm__("" : "+v"(ia), "+v"(a0), "+v"(a1));
                const bool act = (lane >> 4) == (bl & 3) && ia >= 0;
                const int u = act ? ia : n;
                const FT w2 = (FT)2 * (FT)weight_of_words<WT>(a0, a1);
                const FT du = D[u];
                const int xu = x[u] != 0;
                // D_u - 2 w s_k s_u (s_k s_u = 1 when the spins are equal)
                const FT nd = du - (xu == xk ? w2 : -w2);
                D[u] = act ? nd : du;
            } else {
                const int e0 = __builtin_amdgcn_readlane(e0l, bl);
                const int sk = xk ? -1 : 1;
                for (int j = lane; j < deg; j += 64) {
                    const int u = g.ai[e0 + j];
                    const WT w = g.aw[e0 + j];
                    const int su = x[u] ? -1 : 1;
                    D[u] -= (FT)(2 * sk * su) * (FT)w;
                }
            }
            __builtin_amdgcn_wave_barrier();
            // every lane stores the same values at the same addresses
            D[k] = -dd;
            x[k] = (uint8_t)(xk ^ 1);
            age[k] = 0;
            xl = lane == bl ? (uint32_t)(xk ^ 1) : xl;
            __builtin_amdgcn_wave_barrier();
            d = valid ? D[v] : (FT)0;
            mask = __ballot(valid && d > 0 && lane > bl);
        }
    }
    __syncthreads();
    for (int u = lane; u < n; u += 64) {
        const uint8_t xv = x[u];
        hg[u] = xv ? -D[u] : D[u];
        r.x[u] = xv;
        r.age[u] = age[u] < 255 ? (uint8_t)(age[u] + 1) : age[u];
    }
    if (lane == 0) {
        r.out->accepts = flips;
        r.out->gain_sum = gsum;
        r.out->evals = n;
        r.out->gain = gmin;  // the smallest accepted gain (+inf: none)
        r.out->k = -1;
        r.out->l = -1;
    }
}

// One first-improvement pass in score order: nodes are visited in perm order
// and flipped iff their exact gain is positive at that moment.  64 consecutive
// candidates are evaluated at once (one per lane); a ballot finds the first
// positive one, it is applied, and the later lanes re-evaluate.
template <typename WT, typename FT>
__device__ void learned_pass_plain(const BInstRaw& r, const GView<WT>& g, int use_lds, float min_score,
                                   unsigned char* lds);

template <typename WT, typename FT>
__global__ __launch_bounds__(64) void kb_learned_pass(const BInstRaw* insts, const int32_t* list, int use_lds,
                                                     float min_score, LpGate* gate, int pass) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int i = list[blockIdx.x];
    if (gate && !gate[i].active) return;  // wave-uniform: this instance takes no more passes
    const BInstRaw& r = insts[i];
    const GView<WT> g = gview_of<WT>(r);
    if (use_lds == 2 && r.use_adj)
        learned_pass_staged<WT, FT>(r, g, min_score, lds);
    else
        learned_pass_plain<WT, FT>(r, g, use_lds, min_score, lds);
    if (gate && threadIdx.x == 0) {
        // the record lane 0 has just written; the same bookkeeping as the
        // host's (scale is 1 or 2, so the product is exact)
        LpGate& gt = gate[i];
        const PhaseOut o = *r.out;
        gt.outs[pass] = o;
        gt.energy = gt.energy + gt.scale * o.gain_sum;
        gt.active = (o.accepts > 0 && gt.energy < gt.target) ? 1 : 0;
    }
}

template <typename WT, typename FT>
__device__ void learned_pass_plain(const BInstRaw& r, const GView<WT>& g, int use_lds, float min_score,
                                   unsigned char* lds) {
    const int lane = threadIdx.x;
    const int64_t n = r.n;
    FT* hg = (FT*)r.h;
    FT* D = use_lds ? (FT*)lds : hg;  // gains (LDS) or the field itself (global)
    uint8_t* x = use_lds ? lds + (((size_t)n * sizeof(FT) + 15) & ~(size_t)15) : r.x;
    if (use_lds) {
        for (int64_t v = lane; v < n; v += 64) {
            const uint8_t xv = r.x[v];
            x[v] = xv;
            D[v] = xv ? -hg[v] : hg[v];
        }
    }
    __syncthreads();
    auto gain = [&](int64_t v) -> FT { return use_lds ? D[v] : (x[v] ? -hg[v] : hg[v]); };
    int64_t flips = 0;
    double gsum = 0.0, gmin = INFINITY;
    for (int64_t c0 = 0; c0 < n; c0 += 64) {
        const int64_t v = c0 + lane < n ? r.perm[c0 + lane] : 0;
        const bool valid = c0 + lane < n && r.score[v] > min_score;
        FT d = valid ? gain(v) : (FT)0;
        uint64_t mask = __ballot(valid && d > 0);
        while (mask) {
            const int bl = __builtin_ctzll(mask);
            const int64_t k = __shfl(v, bl, 64);
            const FT dd = __shfl(d, bl, 64);
            gsum += (double)dd;
            gmin = fmin(gmin, (double)dd);
            ++flips;
            const int sk = x[k] ? -1 : 1;
            __syncthreads();
            if (use_lds) {
                for (int32_t e = g.ap[k] + lane; e < g.ap[k + 1]; e += 64) {
                    const int32_t u = g.ai[e];
                    const int su = x[u] ? -1 : 1;
                    D[u] -= (FT)(2 * sk * su) * (FT)g.aw[e];
                }
                __syncthreads();
                if (lane == 0) {
                    D[k] = -D[k];
                    x[k] ^= 1u;
                }
            } else {
                flip_update(g, x, hg, k, sk, lane, 64);
            }
            if (lane == 0) r.age[k] = 0;
            __syncthreads();
            d = valid ? gain(v) : (FT)0;
            mask = __ballot(valid && d > 0 && lane > bl);
        }
    }
    __syncthreads();
    for (int64_t u = lane; u < n; u += 64) {
        if (use_lds) {
            const uint8_t xv = x[u];
            hg[u] = xv ? -D[u] : D[u];
            r.x[u] = xv;
        }
        if (r.age[u] < 255) r.age[u]++;
    }
    if (lane == 0) {
        r.out->accepts = flips;
        r.out->gain_sum = gsum;
        r.out->evals = n;
        r.out->gain = gmin;  // the smallest accepted gain (+inf: none)
        r.out->k = -1;
        r.out->l = -1;
    }
}

template <typename WT>
static void b_features_impl(const BInstRaw* d, const int32_t* list, int64_t L, int64_t max_n, hipStream_t s) {
    using FT = typename FieldOf<WT>::T;
    hipLaunchKernelGGL((kb_features<WT, FT>), dim3((unsigned)((max_n + 255) / 256), (unsigned)L), dim3(256), 0, s, d,
                       list);
}

template <typename WT>
static void b_learned_impl(const BInstRaw* d, const int32_t* list, int64_t L, int64_t max_n, float min_score,
                           LpGate* gate, int pass, hipStream_t s) {
    using FT = typename FieldOf<WT>::T;
    // 2: the staged pass (sparse graphs), 1: gains and spins in LDS, 0: global
    size_t lds = (((size_t)max_n * sizeof(FT) + 15) & ~(size_t)15) + (size_t)max_n;
    static const size_t limit = lds_limit((const void*)kb_learned_pass<WT, FT>);
    const size_t staged = pass_lds_layout<WT, FT>(max_n).total;
    int use_lds = lds <= limit ? 1 : 0;
    if (staged <= limit && descent_variant() != 8) {
        use_lds = 2;
        lds = staged;
    }
    hipLaunchKernelGGL((kb_learned_pass<WT, FT>), dim3((unsigned)L), dim3(64), use_lds ? lds : 0, s, d, list, use_lds,
                       min_score, gate, pass);
}

void launch_b_features(int kind, const BInstRaw* d, const int32_t* list, int64_t L, int64_t max_n, hipStream_t s) {
    PQI_KIND_DISPATCH(kind, b_features_impl, d, list, L, max_n, s);
}
void launch_b_order(const BInstRaw* d, const int32_t* list, int64_t L, hipStream_t s) {
    hipLaunchKernelGGL(kb_order, dim3((unsigned)L), dim3(256), 0, s, d, list);
}
void launch_b_learned_pass(int kind, const BInstRaw* d, const int32_t* list, int64_t L, int64_t max_n,
                           float min_score, LpGate* gate, int pass, hipStream_t s) {
    PQI_KIND_DISPATCH(kind, b_learned_impl, d, list, L, max_n, min_score, gate, pass, s);
}

}  // namespace pqi
