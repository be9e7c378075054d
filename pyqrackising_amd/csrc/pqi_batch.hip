// Batched candidate energies over many int8 instances of uneven size in one
// persistent launch, captured in a hipGraph, with tile variants chosen by an
// on-device autotuner.
//
// Each instance i has W_i (n_pad_i x n_pad_i int8) and B candidates packed in
// its own slice of one bits buffer.  The work is a list of 128 x 128 output
// tiles (instance, candidate block, column block), sorted by contraction length
// so that a static round-robin over a persistent grid is a longest-first
// schedule.  Every tile runs the same MFMA tile as the single-instance kernel
// (pqi_tile_i8.h); row totals are exact int64 atomics.  A second kernel picks
// each instance's first best candidate (lowest q = s^T W s, i.e. highest cut and
// highest energy), replacing the host argmax of pyqrackising/maxcut_tfim.py:275.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <string>
#include <vector>

#include "pqi_internal.h"
#include "pqi_tile_i8.h"

namespace pqi {

struct BatchInst {
    const int8_t* W;
    int64_t n_pad;
    const uint32_t* bits;
    int64_t wpr;
    int64_t B;
    unsigned long long* q;
};

struct BatchTile {
    int32_t inst, mb, jb, pad;
};

template <int BK>
__global__ __launch_bounds__(256) void k_batch_grouped(const BatchInst* __restrict__ insts,
                                                       const BatchTile* __restrict__ tiles, int64_t ntiles) {
    __shared__ __align__(16) int8_t As[TILE_BM * (BK + 16)];
    __shared__ __align__(16) int8_t Bs[TILE_BN * (BK + 16)];
    __shared__ uint32_t Sb[TILE_BM * (TILE_BN / 32)];
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const BatchTile tl = tiles[t];
        const BatchInst in = insts[tl.inst];
        tile_i8<BK>(in.W, in.n_pad, in.n_pad, in.bits, in.wpr, in.B, (int64_t)tl.mb * TILE_BM,
                    (int64_t)tl.jb * TILE_BN, in.q, As, Bs, Sb);
    }
}

// Wide tiles (256 candidates x 256 columns, pqi_tile_i8.h), persistent over a
// tile list laid out for the XCDs: position p is run by a work group of XCD
// p % 8, and all tiles of one instance sit at positions of one residue, its
// candidate tiles of a column block next to each other, so they share that
// block of W and the instance's candidate bits through one L2.
// Empty positions (inst < 0) pad the per-XCD queues to one length.
template <int BK, int NW>
__global__ __launch_bounds__(64 * NW) void k_batch_wide(const BatchInst* __restrict__ insts,
                                                        const BatchTile* __restrict__ tiles, int64_t ntiles) {
    extern __shared__ __align__(16) unsigned char lds[];
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const BatchTile tl = tiles[t];
        if (tl.inst < 0) continue;
        const BatchInst in = insts[tl.inst];
        tile_i8_wide<BK, NW>(in.W, in.n_pad, 0, in.n_pad, in.bits, in.wpr, in.B, (int64_t)tl.mb * WIDE_BM,
                         (int64_t)tl.jb * WIDE_BN, in.q, lds);
    }
}

__global__ void k_zero_q(unsigned long long* q, int64_t count) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) q[i] = 0ull;
}

// per instance: first index of the smallest q
__global__ __launch_bounds__(256) void k_seg_best(const BatchInst* __restrict__ insts, int64_t* best,
                                                  long long* best_q) {
    __shared__ long long sv[4];
    __shared__ int64_t sk[4];
    const BatchInst in = insts[blockIdx.x];
    long long bv = LLONG_MAX;
    int64_t bk = LLONG_MAX;
    for (int64_t b = threadIdx.x; b < in.B; b += blockDim.x) {
        const long long v = (long long)in.q[b];
        if (v < bv) { bv = v; bk = b; }
    }
    for (int m = 32; m > 0; m >>= 1) {
        const long long ov = ((long long)__shfl_xor((int)(bv >> 32), m, 64) << 32) |
                             (unsigned)__shfl_xor((int)(bv & 0xffffffff), m, 64);
        const int64_t ok = ((int64_t)__shfl_xor((int)(bk >> 32), m, 64) << 32) |
                           (unsigned)__shfl_xor((int)(bk & 0xffffffff), m, 64);
        if (ov < bv || (ov == bv && ok < bk)) { bv = ov; bk = ok; }
    }
    if ((threadIdx.x & 63) == 0) {
        sv[threadIdx.x >> 6] = bv;
        sk[threadIdx.x >> 6] = bk;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        bv = sv[0];
        bk = sk[0];
        for (int i = 1; i < 4; ++i)
            if (sv[i] < bv || (sv[i] == bv && sk[i] < bk)) { bv = sv[i]; bk = sk[i]; }
        best[blockIdx.x] = bk;
        best_q[blockIdx.x] = bv;
    }
}

// tile variants: 0-5 the 128 x 128 tile, contraction depth x persistent grid
// size (multiples of the CU count); 6-7 the wide tile on 4 waves (one work
// group per CU) with depth 128 or 64; 8-9 the wide tile on 8 waves
static const int kVariantBK[] = {64, 64, 64, 128, 128, 128, 128, 64, 128, 64};
static const int kVariantGrid[] = {1, 2, 4, 1, 2, 4, 1, 1, 1, 1};
static const int kVariantWide[] = {0, 0, 0, 0, 0, 0, 4, 4, 8, 8};
static const int kVariants = 10;

}  // namespace pqi

struct pqi_batch {
    pqi_ctx* ctx = nullptr;
    int64_t count = 0, cands = 0;
    std::vector<pqi_graph*> graphs;
    std::vector<int64_t> bits_off;  // word offset of each instance's slice
    std::vector<int64_t> q_off;
    int64_t total_words = 0, total_q = 0, ntiles = 0;
    pqi::BatchInst* d_insts = nullptr;
    pqi::BatchTile* d_tiles = nullptr;
    pqi::BatchTile* d_wtiles = nullptr;  // wide-tile list (nullptr: wide tiles not usable)
    int64_t nwtiles = 0;
    uint32_t* d_bits = nullptr;
    unsigned long long* d_q = nullptr;
    int64_t* d_best = nullptr;
    long long* d_bestq = nullptr;
    int variant = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int64_t device_bytes = 0;
};

namespace pqi {

static void batch_release(pqi_batch* b) {
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    void* ptrs[] = {b->d_insts, b->d_tiles, b->d_wtiles, b->d_bits, b->d_q, b->d_best, b->d_bestq};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete b;
}

static void enqueue_eval(pqi_batch* b, int variant, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_q, dim3((unsigned)((b->total_q + 255) / 256)), dim3(256), 0, s, b->d_q, b->total_q);
    if (kVariantWide[variant] && !b->d_wtiles) variant = 3;  // wide tiles not usable: the nearest plain variant
    const int grid = std::max(1, b->ctx->num_cus) * kVariantGrid[variant];
    if (kVariantWide[variant]) {
        static const bool lds_ok = [] {
            bool ok = true;
            ok &= hipFuncSetAttribute((const void*)k_batch_wide<128, 4>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)wide_lds_bytes<128>()) == hipSuccess;
            ok &= hipFuncSetAttribute((const void*)k_batch_wide<128, 8>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)wide_lds_bytes<128>()) == hipSuccess;
            return ok;
        }();
        const int nw = kVariantWide[variant];
        const int bk = lds_ok ? kVariantBK[variant] : 64;  // the 64-deep wide tile fits the default LDS limit
        const unsigned g = (unsigned)std::min<int64_t>(grid, std::max<int64_t>(1, b->nwtiles));
        if (bk == 128 && nw == 8)
            hipLaunchKernelGGL((k_batch_wide<128, 8>), dim3(g), dim3(512), wide_lds_bytes<128>(), s, b->d_insts,
                               b->d_wtiles, b->nwtiles);
        else if (bk == 128)
            hipLaunchKernelGGL((k_batch_wide<128, 4>), dim3(g), dim3(256), wide_lds_bytes<128>(), s, b->d_insts,
                               b->d_wtiles, b->nwtiles);
        else if (nw == 8)
            hipLaunchKernelGGL((k_batch_wide<64, 8>), dim3(g), dim3(512), wide_lds_bytes<64>(), s, b->d_insts,
                               b->d_wtiles, b->nwtiles);
        else
            hipLaunchKernelGGL((k_batch_wide<64, 4>), dim3(g), dim3(256), wide_lds_bytes<64>(), s, b->d_insts,
                               b->d_wtiles, b->nwtiles);
        hipLaunchKernelGGL(k_seg_best, dim3((unsigned)b->count), dim3(256), 0, s, b->d_insts, b->d_best, b->d_bestq);
        return;
    }
    const unsigned g = (unsigned)std::min<int64_t>(grid, std::max<int64_t>(1, b->ntiles));
    if (kVariantBK[variant] == 64)
        hipLaunchKernelGGL(k_batch_grouped<64>, dim3(g), dim3(256), 0, s, b->d_insts, b->d_tiles, b->ntiles);
    else
        hipLaunchKernelGGL(k_batch_grouped<128>, dim3(g), dim3(256), 0, s, b->d_insts, b->d_tiles, b->ntiles);
    hipLaunchKernelGGL(k_seg_best, dim3((unsigned)b->count), dim3(256), 0, s, b->d_insts, b->d_best, b->d_bestq);
}

static int capture(pqi_batch* b) {
    hipStream_t s = b->ctx->stream;
    if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
    }
    if (b->graph) {
        (void)hipGraphDestroy(b->graph);
        b->graph = nullptr;
    }
    if (hipStreamBeginCapture(s, hipStreamCaptureModeGlobal) != hipSuccess)
        return fail(PQI_ERR_HIP, "hipStreamBeginCapture failed");
    enqueue_eval(b, b->variant, s);
    if (hipStreamEndCapture(s, &b->graph) != hipSuccess) return fail(PQI_ERR_HIP, "hipStreamEndCapture failed");
    if (hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess)
        return fail(PQI_ERR_HIP, "hipGraphInstantiate failed");
    return PQI_OK;
}

}  // namespace pqi

using namespace pqi;

extern "C" {

int pqi_batch_create(pqi_ctx* c, pqi_graph* const* graphs, int64_t count, int64_t cands, pqi_batch** out) {
    if (!c || !graphs || !out) return fail(PQI_ERR_INVALID, "null argument");
    if (count < 1 || cands < 1) return fail(PQI_ERR_INVALID, "count and cands must be >= 1");
    *out = nullptr;
    if (hipSetDevice(c->device) != hipSuccess) return fail(PQI_ERR_HIP, "hipSetDevice failed");
    pqi_batch* b = new pqi_batch();
    b->ctx = c;
    b->count = count;
    b->cands = cands;
    std::vector<BatchInst> hin(count);
    std::vector<BatchTile> tiles;
    std::vector<int64_t> cost;
    for (int64_t i = 0; i < count; ++i) {
        pqi_graph* g = graphs[i];
        if (!g || g->ctx != c || g->kind != PQI_WEIGHTS_I8 || !g->W) {
            batch_release(b);
            return fail(PQI_ERR_UNSUPPORTED, "batched evaluation needs dense int8 graphs of this context");
        }
        b->graphs.push_back(g);
        b->bits_off.push_back(b->total_words);
        b->q_off.push_back(b->total_q);
        const int64_t wpr = g->n_pad >> 5;
        b->total_words += cands * wpr;
        b->total_q += cands;
        for (int64_t mb = 0; mb < (cands + TILE_BM - 1) / TILE_BM; ++mb)
            for (int64_t jb = 0; jb < g->n_pad / TILE_BN; ++jb) {
                tiles.push_back({(int32_t)i, (int32_t)mb, (int32_t)jb, 0});
                cost.push_back(g->n_pad);
            }
    }
    // longest contraction first (stable, so tiles of one instance stay together)
    std::vector<int64_t> order(tiles.size());
    for (size_t t = 0; t < order.size(); ++t) order[t] = (int64_t)t;
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t d) { return cost[a] > cost[d]; });
    std::vector<BatchTile> sorted(tiles.size());
    for (size_t t = 0; t < order.size(); ++t) sorted[t] = tiles[order[t]];
    b->ntiles = (int64_t)sorted.size();
    // wide tiles: each instance's tiles -- for every 256-column block, its
    // ceil(cands/256) candidate tiles in a row -- go to one of 8 XCD queues,
    // instances dealt longest first to the least loaded queue, so that one
    // XCD's L2 serves both an instance's column blocks of W (to its candidate
    // tiles) and its candidate bits (to all its column blocks)
    bool wide_ok = true;
    for (int64_t i = 0; i < count && wide_ok; ++i)
        wide_ok = graphs[i]->max_strength >= 0 && 256 * graphs[i]->max_strength < (int64_t(1) << 31);
    std::vector<BatchTile> wsorted;
    if (wide_ok) {
        std::vector<int64_t> insts(count);
        for (int64_t i = 0; i < count; ++i) insts[i] = i;
        std::stable_sort(insts.begin(), insts.end(),
                         [&](int64_t a, int64_t d) { return graphs[a]->n_pad > graphs[d]->n_pad; });
        const int64_t MBw = (cands + WIDE_BM - 1) / WIDE_BM;
        std::vector<std::vector<BatchTile>> queue(8);
        std::vector<int64_t> load(8, 0);
        for (const int64_t i : insts) {
            const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            const int64_t JBi = (graphs[i]->n_pad + WIDE_BN - 1) / WIDE_BN;
            for (int64_t jb = 0; jb < JBi; ++jb)
                for (int64_t mb = 0; mb < MBw; ++mb) queue[x].push_back({(int32_t)i, (int32_t)mb, (int32_t)jb, 0});
            load[x] += graphs[i]->n_pad * graphs[i]->n_pad * MBw;
        }
        size_t len = 0;
        for (auto& qx : queue) len = std::max(len, qx.size());
        wsorted.assign(len * 8, BatchTile{-1, 0, 0, 0});
        for (int x = 0; x < 8; ++x)
            for (size_t k = 0; k < queue[x].size(); ++k) wsorted[8 * k + x] = queue[x][k];
        b->nwtiles = (int64_t)wsorted.size();
    }
    auto alloc = [&](void** p, size_t bytes) {
        if (hipMalloc(p, std::max<size_t>(bytes, 16)) != hipSuccess) return false;
        b->device_bytes += (int64_t)bytes;
        return true;
    };
    if ((wide_ok && !alloc((void**)&b->d_wtiles, sizeof(BatchTile) * wsorted.size())) ||
        !alloc((void**)&b->d_insts, sizeof(BatchInst) * count) ||
        !alloc((void**)&b->d_tiles, sizeof(BatchTile) * sorted.size()) ||
        !alloc((void**)&b->d_bits, (size_t)b->total_words * 4) || !alloc((void**)&b->d_q, (size_t)b->total_q * 8) ||
        !alloc((void**)&b->d_best, (size_t)count * 8) || !alloc((void**)&b->d_bestq, (size_t)count * 8)) {
        batch_release(b);
        return fail(PQI_ERR_OOM, "device allocation for the batch failed");
    }
    for (int64_t i = 0; i < count; ++i) {
        pqi_graph* g = graphs[i];
        hin[i] = {(const int8_t*)g->W, g->n_pad, b->d_bits + b->bits_off[i], g->n_pad >> 5, cands, b->d_q + b->q_off[i]};
    }
    if (hipMemcpy(b->d_insts, hin.data(), sizeof(BatchInst) * count, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->d_tiles, sorted.data(), sizeof(BatchTile) * sorted.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(b->d_bits, 0, (size_t)b->total_words * 4) != hipSuccess ||
        (wide_ok && hipMemcpy(b->d_wtiles, wsorted.data(), sizeof(BatchTile) * wsorted.size(),
                              hipMemcpyHostToDevice) != hipSuccess)) {
        batch_release(b);
        return fail(PQI_ERR_HIP, "uploading the batch description failed");
    }
    int r = capture(b);
    if (r != PQI_OK) {
        batch_release(b);
        return r;
    }
    *out = b;
    return PQI_OK;
}

int pqi_batch_free(pqi_batch* b) {
    if (!b) return PQI_OK;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    batch_release(b);
    return PQI_OK;
}

int pqi_batch_set_candidates(pqi_batch* b, const uint32_t* bits) {
    if (!b || !bits) return fail(PQI_ERR_INVALID, "null argument");
    if (hipSetDevice(b->ctx->device) != hipSuccess) return fail(PQI_ERR_HIP, "hipSetDevice failed");
    const uint32_t* src = bits;
    for (int64_t i = 0; i < b->count; ++i) {
        pqi_graph* g = b->graphs[i];
        const int64_t nw = (g->n + 31) >> 5, wpr = g->n_pad >> 5;
        const int r = upload_rows(b->ctx, b->d_bits + b->bits_off[i], (size_t)wpr * 4, src, (size_t)nw * 4,
                                  (size_t)b->cands);
        if (r != PQI_OK) return r;
        // the staging buffer is reused by the next instance
        if (hipStreamSynchronize(b->ctx->stream) != hipSuccess) return fail(PQI_ERR_HIP, "upload sync failed");
        src += b->cands * nw;
    }
    if (hipStreamSynchronize(b->ctx->stream) != hipSuccess) return fail(PQI_ERR_HIP, "upload sync failed");
    return PQI_OK;
}

int pqi_batch_eval(pqi_batch* b, int is_spin_glass, int64_t* best_index, double* best_value, double* values) {
    if (!b || !best_index || !best_value) return fail(PQI_ERR_INVALID, "null argument");
    hipStream_t s = b->ctx->stream;
    if (hipSetDevice(b->ctx->device) != hipSuccess) return fail(PQI_ERR_HIP, "hipSetDevice failed");
    if (hipGraphLaunch(b->exec, s) != hipSuccess) return fail(PQI_ERR_HIP, "hipGraphLaunch failed");
    std::vector<long long> bq(b->count);
    std::vector<unsigned long long> q;
    if (hipMemcpyAsync(best_index, b->d_best, b->count * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(bq.data(), b->d_bestq, b->count * 8, hipMemcpyDeviceToHost, s) != hipSuccess)
        return fail(PQI_ERR_HIP, "reading results failed");
    if (values) {
        q.resize(b->total_q);
        if (hipMemcpyAsync(q.data(), b->d_q, b->total_q * 8, hipMemcpyDeviceToHost, s) != hipSuccess)
            return fail(PQI_ERR_HIP, "reading values failed");
    }
    if (hipStreamSynchronize(s) != hipSuccess) return fail(PQI_ERR_HIP, "batch evaluation failed");
    for (int64_t i = 0; i < b->count; ++i) {
        const pqi_graph* g = b->graphs[i];
        auto val = [&](long long qi) {
            return is_spin_glass ? (double)(-qi / 2) : (double)((2 * g->w_up_i - qi) / 4);
        };
        best_value[i] = val(bq[i]);
        if (values)
            for (int64_t t = 0; t < b->cands; ++t) values[b->q_off[i] + t] = val((long long)q[b->q_off[i] + t]);
    }
    return PQI_OK;
}

int pqi_batch_autotune(pqi_batch* b, int reps, int* chosen, double* ms_per_variant, int max_variants) {
    if (!b || !chosen) return fail(PQI_ERR_INVALID, "null argument");
    if (reps < 1) reps = 1;
    hipStream_t s = b->ctx->stream;
    if (hipSetDevice(b->ctx->device) != hipSuccess) return fail(PQI_ERR_HIP, "hipSetDevice failed");
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return fail(PQI_ERR_HIP, "hipEventCreate failed");
    double best_ms = 1e300;
    int best = 0;
    for (int v = 0; v < kVariants; ++v) {
        enqueue_eval(b, v, s);  // warm-up
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < reps; ++r) enqueue_eval(b, v, s);
        (void)hipEventRecord(e1, s);
        if (hipStreamSynchronize(s) != hipSuccess) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            return fail(PQI_ERR_HIP, "autotune launch failed");
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        const double per = ms / reps;
        if (ms_per_variant && v < max_variants) ms_per_variant[v] = per;
        if (per < best_ms) { best_ms = per; best = v; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->variant = best;
    *chosen = best;
    return capture(b);
}

int pqi_batch_set_variant(pqi_batch* b, int variant) {
    if (!b) return fail(PQI_ERR_INVALID, "null argument");
    if (variant < 0 || variant >= kVariants) return fail(PQI_ERR_INVALID, "variant out of range");
    if (hipSetDevice(b->ctx->device) != hipSuccess) return fail(PQI_ERR_HIP, "hipSetDevice failed");
    b->variant = variant;
    return capture(b);
}

int pqi_batch_info(const pqi_batch* b, int64_t* ntiles, int64_t* algorithmic_ops, int64_t* executed_ops,
                   int64_t* weight_bytes, int* variant) {
    if (!b) return fail(PQI_ERR_INVALID, "null argument");
    int64_t ops = 0, exe = 0, wb = 0;
    const bool wide = kVariantWide[b->variant] && b->d_wtiles;
    for (const pqi_graph* g : b->graphs) {
        ops += 2 * b->cands * g->n * g->n;
        if (wide)  // candidate and column blocks of 256, the last ones partly empty
            exe += 2 * ((b->cands + 255) / 256 * 256) * ((g->n_pad + 255) / 256 * 256) * g->n_pad;
        else
            exe += 2 * b->cands * g->n_pad * g->n_pad;
        wb += g->n * g->n;
    }
    if (ntiles) *ntiles = wide ? b->nwtiles : b->ntiles;
    if (algorithmic_ops) *algorithmic_ops = ops;
    if (executed_ops) *executed_ops = exe;
    if (weight_bytes) *weight_bytes = wb;
    if (variant) *variant = b->variant;
    return PQI_OK;
}

}  // extern "C"
