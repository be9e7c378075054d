// Candidate sampler on the GPU: the neighbour-repulsion bit picker that
// draws the candidates maxcut_tfim scores (pyqrackising/maxcut_tfim.py:36-75
// local_repulsion_choice / update_repulsion_choice, maxcut_tfim_util.py:690-724
// sample_mag / bit_pick), followed by the round's scoring on the device.
//
// One wavefront per shot.  The shot's weights (float64, initially 1) live in
// LDS; lane l owns the contiguous chunk [l c, (l+1) c) and keeps its chunk sum.
// A pick draws u, sets p = u * total, finds the chunk whose running prefix
// passes p (a wave scan of the 64 chunk sums), then scans that chunk 64
// entries at a time across the wave: the first node i with p < cumsum(..i) is
// picked, as bit_pick does.  Picking k zeroes weights[k] and multiplies every
// weight by base^(-w_kv); only the neighbours of k change (w = 0 multiplies by
// exactly 1), so the update walks the neighbours (the packed (w << 24) | u
// table for int8 weights, else the adjacency list or the dense row), flags the
// changed chunks in LDS, and their owner lanes recompute their sums in node
// order.  For int8 weights the factors base^(-w) come from a 255-entry table
// built once per work group.  The scans run on DPP row shifts and broadcasts.
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

struct SamplerArgs {
    const void* W;       // dense weights (or nullptr), row stride ld
    int64_t ld, n;
    const int32_t* ap;   // adjacency lists (used when W is nullptr or use_adj)
    const int32_t* ai;
    const void* aw;
    const int32_t* kadj; // int8 weights: fixed-stride table (w << 24) | u, or nullptr
    int64_t kstride;
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

// base^(-w): from the block's table for int8 weights, else pow
__device__ __forceinline__ double repulsion_factor(const SamplerArgs& a, const double* ptab, double w) {
    return a.kind == PQI_WEIGHTS_I8 ? ptab[(int)w + 127] : pow(a.base, -w);
}

// weights[node] = 0; neighbours v: weights[v] *= base^(-w_node,v).  The chunk
// of every changed entry is flagged in dflag (LDS, one int per lane).
__device__ void repulse(const SamplerArgs& a, const double* ptab, double* wt, int* dflag, int64_t node, int c,
                        bool multiply) {
    const int lane = threadIdx.x & 63;
    if (lane == 0) {
        wt[node] = 0.0;
        dflag[node / c] = 1;
    }
    if (!multiply || a.no_repulsion) return;
    if (a.kadj) {
        const int32_t* row = a.kadj + node * a.kstride;
        for (int64_t e = lane; e < a.kstride; e += 64) {
            const int32_t key = row[e];
            if (key != 0) {  // zero pads the row
                const int v = key & 0xFFFFFF;
                wt[v] *= ptab[(key >> 24) + 127];
                dflag[v / c] = 1;
            }
        }
    } else if (a.use_adj) {
        const int32_t e1 = a.ap[node + 1];
        for (int32_t e = a.ap[node] + lane; e < e1; e += 64) {
            const int v = a.ai[e];
            if (v != node) {
                wt[v] *= repulsion_factor(a, ptab, weight_at(a, a.aw, e));
                dflag[v / c] = 1;
            }
        }
    } else {
        for (int64_t v = lane; v < a.n; v += 64) {
            const double w = weight_at(a, a.W, node * a.ld + v);
            if (w != 0.0 && v != node) {
                wt[v] *= repulsion_factor(a, ptab, w);
                dflag[v / c] = 1;
            }
        }
    }
}

// lane l's chunk sum, in node order
__device__ __forceinline__ double chunk_sum(const double* wt, int64_t n, int c, int lane) {
    const int64_t lo = (int64_t)lane * c, hi = lo + c < n ? lo + c : n;
    double s = 0.0;
#pragma unroll 8
    for (int64_t i = lo; i < hi; ++i) s += wt[i];
    return s;
}

// DPP move with 0 in every lane the pattern does not write
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_or_zero(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWS, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWS, 0xF, false);
    return __hiloint2double(hi, lo);
}

// inclusive scan over the 64 lanes, in lane order: row_shr 1, 2, 4, 8 within
// rows of 16, then row_bcast 15 and 31 across rows
__device__ __forceinline__ double wave_scan_d(double v) {
    v += dpp_or_zero<0x111, 0xF>(v);
    v += dpp_or_zero<0x112, 0xF>(v);
    v += dpp_or_zero<0x114, 0xF>(v);
    v += dpp_or_zero<0x118, 0xF>(v);
    v += dpp_or_zero<0x142, 0xA>(v);
    v += dpp_or_zero<0x143, 0xC>(v);
    return v;
}

// the value of a uniform lane
__device__ __forceinline__ double lane_value(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

// bit_pick (maxcut_tfim_util.py:711-724): the first i with p < cumsum(weights[0..i]),
// p = total * u; node 0 if there is none.  The chunk sums are scanned across
// the wave; the owning chunk is then scanned 64 entries at a time by the whole
// wave.  Only chunks and entries with positive weight can be chosen, so a
// rounding difference between the chunk sums and the walk can never return a
// node already picked.
__device__ int64_t pick(const SamplerArgs& a, const double* wt, double mysum, int c, double u) {
    const int lane = threadIdx.x & 63;
    const double pre = wave_scan_d(mysum);
    const double p = lane_value(pre, 63) * u;
    const uint64_t hit = __ballot(p < pre && mysum > 0.0);
    int owner;
    if (hit) {
        owner = __ffsll((unsigned long long)hit) - 1;
    } else {
        // no positive chunk passes p: all weights zero (the reference returns
        // node 0), or p fell in a rounding gap above the last positive chunk
        const uint64_t nzc = __ballot(mysum > 0.0);
        if (!nzc) return 0;
        owner = 63 - __clzll((unsigned long long)nzc);
    }
    double before = lane_value(pre - mysum, owner);
    const int64_t lo = (int64_t)owner * c, hi = lo + c < a.n ? lo + c : a.n;
    int64_t last = -1;
    for (int64_t base = lo; base < hi; base += 64) {
        const int64_t i = base + lane;
        const double v = i < hi ? wt[i] : 0.0;
        const double q = wave_scan_d(v);
        const uint64_t h2 = __ballot(v > 0.0 && p < before + q);
        if (h2) return base + __ffsll((unsigned long long)h2) - 1;
        const uint64_t nz = __ballot(v > 0.0);
        if (nz) last = base + 63 - __clzll((unsigned long long)nz);
        before += lane_value(q, 63);
    }
    return last >= 0 ? last : lo;
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

__global__ __launch_bounds__(256) void k_sample_repulsion(SamplerArgs a) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int64_t n = a.n;
    const int c = (int)((n + 63) / 64);
    double* ptab = (double*)lds;  // base^(-w), w = -127..127
    double* wt = ptab + 256 + (size_t)wid * ((size_t)c * 64);
    int* dflag = (int*)(ptab + 256 + (size_t)(blockDim.x >> 6) * ((size_t)c * 64)) + wid * 64;
    if (a.kind == PQI_WEIGHTS_I8)
        for (int t = threadIdx.x; t < 255; t += blockDim.x) ptab[t] = pow(a.base, -(double)(t - 127));
    __syncthreads();
    const int64_t shot = (int64_t)blockIdx.x * (blockDim.x >> 6) + wid;
    if (shot >= a.shots) return;  // whole wave: no barrier below
    const int64_t s = a.shot0 + shot;

    const int64_t m = sample_mag(a.cum, a.ncum, uniform01(a.seed, (uint64_t)s, 0));
    const int64_t first = s % n;

    for (int64_t i = lane; i < (int64_t)c * 64; i += 64) wt[i] = i < n ? 1.0 : 0.0;
    dflag[lane] = 0;
    if (m != 1) {
        repulse(a, ptab, wt, dflag, first, c, true);
        // picks 1 .. m-2 repulse; the last pick (pick 1 when m < 2) only zeroes
        const int64_t last = m - 1 > 1 ? m - 1 : 1;
        double mysum = chunk_sum(wt, n, c, lane);
        dflag[lane] = 0;
        for (int64_t t = 1; t <= last; ++t) {
            const int64_t node = pick(a, wt, mysum, c, uniform01(a.seed, (uint64_t)s, (uint64_t)t));
            repulse(a, ptab, wt, dflag, node, c, t < last);
            if (dflag[lane]) {  // this lane's chunk changed
                mysum = chunk_sum(wt, n, c, lane);
                dflag[lane] = 0;
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
    return 256 * 8 + (size_t)waves * ((size_t)c * 64 * 8 + 64 * 4);
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
    a.kadj = g.key_adj;
    a.kstride = g.key_adj_stride;
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
