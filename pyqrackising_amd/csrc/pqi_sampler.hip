// Candidate sampler on the GPU: the neighbour-repulsion bit picker that
// draws the candidates maxcut_tfim scores (pyqrackising/maxcut_tfim.py:36-75
// local_repulsion_choice / update_repulsion_choice, maxcut_tfim_util.py:690-724
// sample_mag / bit_pick), followed by the round's scoring on the device.
//
// One wavefront per shot.  The shot's weights (float64, initially 1) live in
// LDS; lane l owns the contiguous chunk [l c, (l+1) c) and keeps its chunk sum.
// A pick draws u, sets p = u * total, finds the lane whose running prefix
// passes p (a wave scan of the 64 chunk sums) and that lane walks its chunk:
// the first node i with p < cumsum(..i) is picked, as bit_pick does.  Picking
// k zeroes weights[k] and multiplies every weight by base^(-w_kv); only the
// neighbours of k change (w = 0 multiplies by exactly 1), so the update walks
// the adjacency list (or the dense row) and only the lanes owning a changed
// weight recompute their chunk sums, in a fixed order.
//
// The random stream is counter-based (a splitmix64 hash of seed, shot and
// draw index), so a run is reproducible; the reference draws from numba's
// unseedable generator, so its streams cannot be matched in any case.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <vector>

#include "pqi_internal.h"

namespace pqi {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// uniform in [0, 1): 53 random bits
__device__ __forceinline__ double uniform01(uint64_t seed, uint64_t shot, uint64_t draw) {
    const uint64_t z = mix64(seed ^ mix64(shot * 0xD1B54A32D192ED03ull + draw));
    return (double)(z >> 11) * 0x1.0p-53;
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

struct SamplerArgs {
    const void* W;       // dense weights (or nullptr), row stride ld
    int64_t ld, n;
    const int32_t* ap;   // adjacency lists (used when W is nullptr or use_adj)
    const int32_t* ai;
    const void* aw;
    int use_adj, kind;   // PQI_WEIGHTS_*
    const double* cum;   // Hamming-weight CDF (ncum entries, last = 1)
    int64_t ncum;
    double base;         // repulsion base
    int no_repulsion;
    uint64_t seed;
    int64_t shot0;       // index of the first shot (the reference's s)
    int64_t shots;
    uint32_t* bits;      // shots x wpr packed output, LSB-first
    int64_t wpr;
};

__device__ __forceinline__ double weight_at(const SamplerArgs& a, const void* w, int64_t e) {
    switch (a.kind) {
        case PQI_WEIGHTS_I8: return (double)((const int8_t*)w)[e];
        case PQI_WEIGHTS_F32: return (double)((const float*)w)[e];
        default: return ((const double*)w)[e];
    }
}

// weights[node] = 0; neighbours v: weights[v] *= base^(-w_node,v).  Every
// lane that owns a changed entry is flagged in `dirty` (LDS, one per lane).
__device__ void repulse(const SamplerArgs& a, double* wt, int64_t node, int c, int* dirty, bool multiply) {
    const int lane = threadIdx.x & 63;
    if (lane == 0) {
        wt[node] = 0.0;
        dirty[node / c] = 1;
    }
    if (!multiply || a.no_repulsion) return;
    if (a.use_adj) {
        for (int32_t e = a.ap[node] + lane; e < a.ap[node + 1]; e += 64) {
            const int v = a.ai[e];
            const double w = weight_at(a, a.aw, e);
            if (v != node) {
                wt[v] *= pow(a.base, -w);
                dirty[v / c] = 1;
            }
        }
    } else {
        for (int64_t v = lane; v < a.n; v += 64) {
            const double w = weight_at(a, a.W, node * a.ld + v);
            if (w != 0.0 && v != node) {
                wt[v] *= pow(a.base, -w);
                dirty[v / c] = 1;
            }
        }
    }
}

// sample_mag (maxcut_tfim_util.py:690-708), the reference's own search step
// by step, which matters because the reference's prior is not always
// monotone: mid = (left + right) >> 1 (floor), a negative index wraps once
// (numba), stop at cum[m] >= p and (m == 0 or cum[m-1] < p).  On an input
// where that loop would not end, the lower bound of p is taken after 4096
// steps.
__device__ int64_t sample_mag(const double* cum, int64_t ncum, double p) {
    auto at = [&](int64_t i) { return cum[i < 0 ? i + ncum : i]; };
    int64_t left = 0, right = ncum - 1;
    for (int it = 0; it < 4096; ++it) {
        const int64_t m = (left + right) >> 1;
        if (m < -ncum || m >= ncum || m - 1 < -ncum) break;
        if (at(m) >= p && (m == 0 || at(m - 1) < p)) return m;
        if (at(m) < p) left = m + 1;
        else right = m - 1;
    }
    int64_t lo = 0, hi = ncum - 1;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (cum[mid] >= p) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// bit_pick (maxcut_tfim_util.py:711-724): first i with p < cumsum(weights[0..i]),
// p = total * u; node 0 if none.
__device__ int64_t pick(const SamplerArgs& a, const double* wt, double mysum, int c, double u) {
    const int lane = threadIdx.x & 63;
    // inclusive prefix of the chunk sums over the lanes, in lane order
    double pre = mysum;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const double o = __shfl_up(pre, d, 64);
        if (lane >= d) pre += o;
    }
    const double total = __shfl(pre, 63, 64);
    const double p = total * u;
    const double before = pre - mysum;
    const uint64_t hit = __ballot(p < pre);
    if (hit == 0) return 0;
    const int owner = __ffsll((unsigned long long)hit) - 1;
    int64_t node = 0;
    if (lane == owner) {
        double cum = before;
        const int64_t lo = (int64_t)lane * c, hi = lo + c < a.n ? lo + c : a.n;
        node = hi - 1;
        for (int64_t i = lo; i < hi; ++i) {
            cum += wt[i];
            if (p < cum) {
                node = i;
                break;
            }
        }
    }
    return __shfl(node, owner, 64);
}

__global__ __launch_bounds__(256) void k_sample_repulsion(SamplerArgs a) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int64_t n = a.n;
    const int c = (int)((n + 63) / 64);
    double* wt = (double*)lds + (size_t)wid * ((size_t)c * 64);
    int* dirty = (int*)((double*)lds + (size_t)(blockDim.x >> 6) * ((size_t)c * 64)) + wid * 64;
    const int64_t shot = (int64_t)blockIdx.x * (blockDim.x >> 6) + wid;
    if (shot >= a.shots) return;  // whole wave: no barrier below
    const int64_t s = a.shot0 + shot;

    const int64_t m = sample_mag(a.cum, a.ncum, uniform01(a.seed, (uint64_t)s, 0));
    const int64_t first = s % n;

    for (int64_t i = lane; i < (int64_t)c * 64; i += 64) wt[i] = i < n ? 1.0 : 0.0;
    dirty[lane] = 0;
    if (m != 1) {
        repulse(a, wt, first, c, dirty, true);
        // picks 1 .. m-2 repulse; the last pick (pick 1 when m < 2) only zeroes
        const int64_t last = m - 1 > 1 ? m - 1 : 1;
        double mysum = 0.0;
        for (int64_t i = (int64_t)lane * c; i < (int64_t)(lane + 1) * c && i < n; ++i) mysum += wt[i];
        dirty[lane] = 0;
        for (int64_t t = 1; t <= last; ++t) {
            const int64_t node = pick(a, wt, mysum, c, uniform01(a.seed, (uint64_t)s, (uint64_t)t));
            repulse(a, wt, node, c, dirty, t < last);
            if (dirty[lane]) {
                mysum = 0.0;
                for (int64_t i = (int64_t)lane * c; i < (int64_t)(lane + 1) * c && i < n; ++i) mysum += wt[i];
                dirty[lane] = 0;
            }
        }
    }
    // selected = weight exactly 0 (m == 1: the first node alone)
    uint32_t* row = a.bits + shot * a.wpr;
    for (int64_t w = lane; w < a.wpr; w += 64) {
        uint32_t word = 0;
        for (int q = 0; q < 32; ++q) {
            const int64_t v = w * 32 + q;
            if (v >= n) break;
            const bool sel = m == 1 ? v == first : wt[v] == 0.0;
            if (sel) word |= 1u << q;
        }
        row[w] = word;
    }
}

size_t sampler_lds_bytes(int64_t n, int waves) {
    const int64_t c = (n + 63) / 64;
    return (size_t)waves * ((size_t)c * 64 * 8 + 64 * 4);
}

int launch_sampler(const Graph& g, const double* cum, int64_t ncum, double base, uint64_t seed, int64_t shot0,
                   int64_t shots, uint32_t* bits, int64_t wpr, hipStream_t s) {
    SamplerArgs a;
    a.W = g.W;
    a.ld = g.n_pad;
    a.n = g.n;
    a.ap = g.adj_ptr;
    a.ai = g.adj_idx;
    a.aw = g.adj_w;
    a.use_adj = (g.use_adj || g.W == nullptr) ? 1 : 0;
    a.kind = g.kind;
    a.cum = cum;
    a.ncum = ncum;
    a.base = base;
    a.no_repulsion = std::fabs(1.0 - base) <= 0x1.0p-23 ? 1 : 0;  // maxcut_tfim.py:41 (epsilon)
    a.seed = seed;
    a.shot0 = shot0;
    a.shots = shots;
    a.bits = bits;
    a.wpr = wpr;
    static const size_t limit = [] {
        const int want = 160 * 1024;
        if (hipFuncSetAttribute((const void*)k_sample_repulsion, hipFuncAttributeMaxDynamicSharedMemorySize, want) ==
            hipSuccess)
            return (size_t)want;
        (void)hipGetLastError();
        return (size_t)64 * 1024;
    }();
    int waves = 4;
    while (waves > 1 && sampler_lds_bytes(g.n, waves) > limit) waves >>= 1;
    if (sampler_lds_bytes(g.n, waves) > limit) return fail(PQI_ERR_UNSUPPORTED, "graph too large for the sampler");
    const int64_t blocks = (shots + waves - 1) / waves;
    hipLaunchKernelGGL(k_sample_repulsion, dim3((unsigned)blocks), dim3(64 * waves), sampler_lds_bytes(g.n, waves),
                       s, a);
    return PQI_OK;
}

}  // namespace pqi
