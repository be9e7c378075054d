// libpqi C ABI: contexts, graph upload, and the host side of the local search.
//
// The host code here takes the place of the reference's OpenCL plumbing
// (pyqrackising/maxcut_tfim_util.py:9-301) and of the Python loops that drive
// each phase (pyqrackising/spin_glass_solver.py:603-727).  Decisions between
// phases are made on the host from a few bytes read back per phase; all
// arrays stay resident in HBM for the whole search.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "pqi_internal.h"

namespace pqi {

static thread_local std::string g_error;

void set_error(const std::string& msg) { g_error = msg; }

int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

#define HIPCHK(expr)                                                                             \
    do {                                                                                         \
        hipError_t _e = (expr);                                                                  \
        if (_e != hipSuccess)                                                                    \
            return fail(_e == hipErrorOutOfMemory ? PQI_ERR_OOM : PQI_ERR_HIP,                   \
                        std::string(#expr) + ": " + hipGetErrorString(_e));                      \
    } while (0)

#define CHK(expr)                       \
    do {                                \
        int _r = (expr);                \
        if (_r != PQI_OK) return _r;    \
    } while (0)

// Scratch slots of a context.
enum Slot {
    S_X, S_H, S_RX, S_RH, S_FIX, S_WX, S_WH, S_WE, S_WEV, S_SEL, S_SELD, S_DELTA, S_Q, S_BITS, S_BATCH,
    S_OUT, S_DBL, S_SCORE, S_HB, S_QB, S_RANK, S_CTL, S_OUT2, S_CUM, S_STAGE, S_BARENA, S_BARENA2, S_COUNT
};

static int ensure(pqi_ctx* c, int slot, size_t bytes, void** out) {
    if ((int)c->bufs.size() < S_COUNT) c->bufs.resize(S_COUNT);
    Buffer& b = c->bufs[slot];
    if (bytes == 0) bytes = 16;
    if (b.bytes < bytes) {
        // queued work may still use the old buffer (a local search returns with
        // its last, idle iteration still on the stream; the second batched
        // stream's arena likewise)
        if (b.p) HIPCHK(slot == S_BARENA2 ? hipStreamSynchronize(c->stream2) : hipStreamSynchronize(c->stream));
        if (b.p) HIPCHK(hipFree(b.p));
        b.p = nullptr;
        b.bytes = 0;
        size_t want = std::max(bytes, b.bytes + b.bytes / 2);
        HIPCHK(hipMalloc(&b.p, want));
        b.bytes = want;
    }
    *out = b.p;
    return PQI_OK;
}

int batch_arena(pqi_ctx* c, int set, size_t bytes, void** out) { return ensure(c, set ? S_BARENA2 : S_BARENA, bytes, out); }

// dst row r = src row r (width bytes) followed by zeros up to dpitch
__global__ void k_restride_words(uint32_t* __restrict__ dst, int64_t dwords, const uint32_t* __restrict__ src,
                                 int64_t swords, int64_t rows) {
    const int64_t total = dwords * rows;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = t / dwords, w = t - r * dwords;
        dst[t] = w < swords ? src[r * swords + w] : 0u;
    }
}

__global__ void k_restride_bytes(uint8_t* __restrict__ dst, int64_t dpitch, const uint8_t* __restrict__ src,
                                 int64_t width, int64_t rows) {
    const int64_t total = dpitch * rows;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = t / dpitch, b = t - r * dpitch;
        dst[t] = b < width ? src[r * width + b] : (uint8_t)0;
    }
}

void launch_restride_words(uint32_t* dst, int64_t dwords, const uint32_t* src, int64_t swords, int64_t rows,
                           hipStream_t s) {
    const int64_t total = dwords * rows;
    const unsigned blocks = (unsigned)std::min<int64_t>((total + 255) / 256, 65536);
    if (total > 0)
        hipLaunchKernelGGL(k_restride_words, dim3(blocks), dim3(256), 0, s, dst, dwords, src, swords, rows);
}

int upload_rows(pqi_ctx* c, void* dst, size_t dpitch, const void* src, size_t width, size_t rows) {
    if (rows == 0) return PQI_OK;
    if (dpitch == width) {
        HIPCHK(hipMemcpyAsync(dst, src, width * rows, hipMemcpyHostToDevice, c->stream));
        return PQI_OK;
    }
    void* stage;
    CHK(ensure(c, S_STAGE, width * rows, &stage));
    HIPCHK(hipMemcpyAsync(stage, src, width * rows, hipMemcpyHostToDevice, c->stream));
    const int64_t total = (int64_t)(dpitch * rows);
    const unsigned blocks = (unsigned)std::min<int64_t>((total + 255) / 256, 65536);
    if (dpitch % 4 == 0 && width % 4 == 0 && ((uintptr_t)dst & 3) == 0)
        hipLaunchKernelGGL(k_restride_words, dim3(blocks), dim3(256), 0, c->stream, (uint32_t*)dst,
                           (int64_t)(dpitch / 4), (const uint32_t*)stage, (int64_t)(width / 4), (int64_t)rows);
    else
        hipLaunchKernelGGL(k_restride_bytes, dim3(blocks), dim3(256), 0, c->stream, (uint8_t*)dst, (int64_t)dpitch,
                           (const uint8_t*)stage, (int64_t)width, (int64_t)rows);
    HIPCHK(hipGetLastError());
    return PQI_OK;
}

static int sync(pqi_ctx* c) {
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    return PQI_OK;
}

static int set_device(pqi_ctx* c) {
    HIPCHK(hipSetDevice(c->device));
    return PQI_OK;
}

// ---------------------------------------------------------------------------
// Graph construction (host side)
// ---------------------------------------------------------------------------

template <typename T>
static double as_double(const void* p, int64_t i) {
    return (double)((const T*)p)[i];
}

static double read_w(const void* p, int dtype, int64_t i) {
    switch (dtype) {
        case PQI_DTYPE_I8: return as_double<int8_t>(p, i);
        case PQI_DTYPE_I32: return as_double<int32_t>(p, i);
        case PQI_DTYPE_F32: return as_double<float>(p, i);
        default: return as_double<double>(p, i);
    }
}

static bool valid_dtype(int d) {
    return d == PQI_DTYPE_I8 || d == PQI_DTYPE_I32 || d == PQI_DTYPE_F32 || d == PQI_DTYPE_F64;
}

static int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

// Upload the symmetric adjacency (rows sorted) in the graph's element type.
static int upload_adjacency(pqi_graph* g, const std::vector<int32_t>& ptr, const std::vector<int32_t>& idx,
                            const std::vector<double>& w) {
    const int64_t nnz = (int64_t)idx.size();
    g->nnz = nnz;
    if (nnz > INT32_MAX / 2) return fail(PQI_ERR_UNSUPPORTED, "too many nonzeros for the adjacency lists");
    HIPCHK(hipMalloc(&g->adj_ptr, sizeof(int32_t) * ptr.size()));
    HIPCHK(hipMalloc(&g->adj_idx, sizeof(int32_t) * std::max<int64_t>(1, nnz)));
    size_t es = g->kind == PQI_WEIGHTS_I8 ? 1 : (g->kind == PQI_WEIGHTS_F32 ? 4 : 8);
    HIPCHK(hipMalloc(&g->adj_w, es * std::max<int64_t>(1, nnz)));
    HIPCHK(hipMemcpy(g->adj_ptr, ptr.data(), sizeof(int32_t) * ptr.size(), hipMemcpyHostToDevice));
    if (nnz) {
        HIPCHK(hipMemcpy(g->adj_idx, idx.data(), sizeof(int32_t) * nnz, hipMemcpyHostToDevice));
        if (g->kind == PQI_WEIGHTS_I8) {
            std::vector<int8_t> t(nnz);
            for (int64_t i = 0; i < nnz; ++i) t[i] = (int8_t)w[i];
            HIPCHK(hipMemcpy(g->adj_w, t.data(), nnz, hipMemcpyHostToDevice));
        } else if (g->kind == PQI_WEIGHTS_F32) {
            std::vector<float> t(nnz);
            for (int64_t i = 0; i < nnz; ++i) t[i] = (float)w[i];
            HIPCHK(hipMemcpy(g->adj_w, t.data(), nnz * 4, hipMemcpyHostToDevice));
        } else {
            HIPCHK(hipMemcpy(g->adj_w, w.data(), nnz * 8, hipMemcpyHostToDevice));
        }
    }
    g->device_bytes += (int64_t)(sizeof(int32_t) * (ptr.size() + nnz) + es * nnz);
    int64_t md = 0;
    double ms = 0.0;
    for (size_t i = 0; i + 1 < ptr.size(); ++i) {
        md = std::max<int64_t>(md, ptr[i + 1] - ptr[i]);
        double r = 0.0;
        for (int32_t e = ptr[i]; e < ptr[i + 1]; ++e) r += std::fabs(w[e]);
        ms = std::max(ms, r);
    }
    g->max_degree = md;
    if (g->kind == PQI_WEIGHTS_I8) {
        g->max_strength = (int64_t)ms;
        g->unit_abs = std::all_of(w.begin(), w.end(), [](double v) { return std::fabs(v) == 1.0; });
        int par = ptr.size() > 1 ? (int)((ptr[1] - ptr[0]) & 1) : 0;
        for (size_t i = 0; i + 1 < ptr.size() && par >= 0; ++i)
            if ((int)((ptr[i + 1] - ptr[i]) & 1) != par) par = -1;
        g->degree_parity = par;
        // fixed-stride table for the packed-key descent when padding stays small
        const int64_t n = (int64_t)ptr.size() - 1;
        if (md > 0 && md <= 64 && n < (1 << 24) && n * md <= 4 * nnz) {
            const int64_t words = (n * md + 3) & ~int64_t(3);
            std::vector<int32_t> t(words, 0);
            for (int64_t v = 0; v < n; ++v)
                for (int32_t e = ptr[v]; e < ptr[v + 1]; ++e)
                    t[v * md + (e - ptr[v])] = (int32_t)((uint32_t)(uint8_t)(int8_t)w[e] << 24) | idx[e];
            HIPCHK(hipMalloc(&g->key_adj, words * 4));
            HIPCHK(hipMemcpy(g->key_adj, t.data(), words * 4, hipMemcpyHostToDevice));
            g->key_adj_stride = md;
            g->device_bytes += words * 4;
        }
    }
    return PQI_OK;
}

static int classify(double maxabs, bool all_int, int dtype) {
    if (all_int && maxabs <= 127.0) return PQI_WEIGHTS_I8;
    if (dtype == PQI_DTYPE_F32) return PQI_WEIGHTS_F32;
    return PQI_WEIGHTS_F64;
}

static void ls_graphs_reset(pqi_ctx* c) {
    (void)hipStreamSynchronize(c->stream);  // a replay may still be queued
    for (auto& e : c->ls.seg)
        if (e) {
            (void)hipGraphExecDestroy(e);
            e = nullptr;
        }
    for (hipGraphExec_t* e : {&c->ls.iter, &c->ls.iter2})
        if (*e) {
            (void)hipGraphExecDestroy(*e);
            *e = nullptr;
        }
    c->ls.key.clear();
    c->ls.failed = false;
}

static void free_graph(pqi_graph* g) {
    if (!g) return;
    if (g->W) (void)hipFree(g->W);
    if (g->Wbf) (void)hipFree(g->Wbf);
    if (g->adj_ptr) (void)hipFree(g->adj_ptr);
    if (g->adj_idx) (void)hipFree(g->adj_idx);
    if (g->adj_w) (void)hipFree(g->adj_w);
    if (g->key_adj) (void)hipFree(g->key_adj);
    delete g;
}

}  // namespace pqi

using namespace pqi;

extern "C" {

const char* pqi_last_error(void) { return g_error.c_str(); }

int pqi_version(void) { return 1; }

int pqi_device_count(int* count) {
    if (!count) return fail(PQI_ERR_INVALID, "count is null");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess || c <= 0) {
        *count = 0;
        return fail(PQI_ERR_NO_DEVICE, std::string("no HIP device: ") + hipGetErrorString(e));
    }
    *count = c;
    return PQI_OK;
}

int pqi_create(int device, pqi_ctx** out) {
    if (!out) return fail(PQI_ERR_INVALID, "out is null");
    *out = nullptr;
    int count = 0;
    CHK(pqi_device_count(&count));
    if (device < 0 || device >= count)
        return fail(PQI_ERR_NO_DEVICE, "device index " + std::to_string(device) + " out of range");
    HIPCHK(hipSetDevice(device));
    pqi_ctx* c = new pqi_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(PQI_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
    c->bufs.resize(S_COUNT);
    // the device address of pqi_eval_batch_noalloc's scratch array on this device
    const int ir = noalloc_init(device);
    if (ir != PQI_OK) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return ir;
    }
    *out = c;
    return PQI_OK;
}

int pqi_destroy(pqi_ctx* c) {
    if (!c) return PQI_OK;
    (void)hipSetDevice(c->device);
    if (c->worker) worker_stop(c->worker);
    (void)hipStreamSynchronize(c->stream);
    if (c->stream2) (void)hipStreamSynchronize(c->stream2);
    ls_graphs_reset(c);
    if (c->ls.host_ctl) (void)hipHostFree(c->ls.host_ctl);
    if (c->ls.seq_dev) (void)hipFree(c->ls.seq_dev);
    for (auto& pb : c->pin)
        if (pb.p) (void)hipHostFree(pb.p);
    for (auto& b : c->bufs)
        if (b.p) (void)hipFree(b.p);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->stream2) (void)hipStreamDestroy(c->stream2);
    delete c;
    return PQI_OK;
}

int pqi_graph_dense(pqi_ctx* c, const void* W, int dtype, int64_t n, pqi_graph** out) {
    if (!c || !W || !out) return fail(PQI_ERR_INVALID, "null argument");
    if (!valid_dtype(dtype)) return fail(PQI_ERR_INVALID, "unsupported dtype");
    if (n < 1 || n > (1 << 20)) return fail(PQI_ERR_INVALID, "n out of range");
    *out = nullptr;
    CHK(set_device(c));
    // pass 1 (64 x 64 tiles for cache locality): diagonal, finiteness, symmetry,
    // value class, nonzero count and W_up = sum_{i<j} w_ij (long double)
    double maxabs = 0.0;
    bool all_int = true;
    int64_t nnz = 0;
    long double w_up_l = 0.0L;
    std::vector<double> strength(n, 0.0);  // row sums of |w|
    const int64_t TB = 64;
    for (int64_t i0 = 0; i0 < n; i0 += TB)
        for (int64_t j0 = i0; j0 < n; j0 += TB)
            for (int64_t i = i0; i < std::min(n, i0 + TB); ++i)
                for (int64_t j = std::max(j0, i); j < std::min(n, j0 + TB); ++j) {
                    const double a = read_w(W, dtype, i * n + j);
                    if (j == i) {
                        if (a != 0.0)
                            return fail(PQI_ERR_INVALID, "W has a nonzero diagonal entry at " + std::to_string(i));
                        continue;
                    }
                    if (std::isnan(a) || std::isinf(a)) return fail(PQI_ERR_INVALID, "W has a non-finite entry");
                    if (a != read_w(W, dtype, j * n + i)) return fail(PQI_ERR_INVALID, "W is not symmetric");
                    if (a != 0.0) {
                        nnz += 2;
                        maxabs = std::max(maxabs, std::fabs(a));
                        if (all_int && a != std::nearbyint(a)) all_int = false;
                        w_up_l += a;
                        strength[i] += std::fabs(a);
                        strength[j] += std::fabs(a);
                    }
                }
    pqi_graph* g = new pqi_graph();
    g->ctx = c;
    g->n = n;
    const double max_strength = n ? *std::max_element(strength.begin(), strength.end()) : 0.0;
    // int8 tiles sum up to 128 columns of |P| <= max strength in int32
    g->kind = classify(maxabs, all_int && 128.0 * max_strength < 2147483648.0, dtype);
    g->n_pad = round_up(n, 128);
    g->w_up = (double)w_up_l;
    g->w_up_i = g->kind == PQI_WEIGHTS_I8 ? (int64_t)std::llround((double)w_up_l) : 0;
    g->nnz = nnz;
    g->use_adj = nnz <= n * n / 8;
    if (g->kind == PQI_WEIGHTS_I8) g->max_strength = (int64_t)max_strength;
    std::vector<double>().swap(strength);
    // dense copy, zero padded, in the storage type (one pitched copy)
    const size_t es = g->kind == PQI_WEIGHTS_I8 ? 1 : (g->kind == PQI_WEIGHTS_F32 ? 4 : 8);
    const size_t bytes = (size_t)g->n_pad * (size_t)g->n_pad * es;
    auto bail = [&](int code, const std::string& m) {
        free_graph(g);
        return fail(code, m);
    };
    if (hipMalloc(&g->W, bytes) != hipSuccess) return bail(PQI_ERR_OOM, "device allocation of W failed");
    g->device_bytes += (int64_t)bytes;
    if (hipMemset(g->W, 0, bytes) != hipSuccess) return bail(PQI_ERR_HIP, "clearing W failed");
    const bool same = (es == 1 && dtype == PQI_DTYPE_I8) || (es == 4 && dtype == PQI_DTYPE_F32) ||
                      (es == 8 && dtype == PQI_DTYPE_F64);
    std::vector<unsigned char> conv;
    const void* src = W;
    if (!same) {
        conv.resize((size_t)n * n * es);
        for (int64_t t = 0; t < n * n; ++t) {
            const double a = read_w(W, dtype, t);
            if (es == 1) ((int8_t*)conv.data())[t] = (int8_t)a;
            else if (es == 4) ((float*)conv.data())[t] = (float)a;
            else ((double*)conv.data())[t] = a;
        }
        src = conv.data();
    }
    if (upload_rows(c, g->W, (size_t)g->n_pad * es, src, (size_t)n * es, (size_t)n) != PQI_OK ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return bail(PQI_ERR_HIP, "uploading W failed");
    std::vector<unsigned char>().swap(conv);
    if (g->kind == PQI_WEIGHTS_F32) {
        const size_t bb = (size_t)g->n_pad * (size_t)g->n_pad * 4;  // two bf16 copies
        if (hipMalloc(&g->Wbf, bb) != hipSuccess) return bail(PQI_ERR_OOM, "device allocation of bf16 W failed");
        g->device_bytes += (int64_t)bb;
        launch_split_bf16((const float*)g->W, g->n_pad, (uint16_t*)g->Wbf, c->stream);
        if (hipStreamSynchronize(c->stream) != hipSuccess) return bail(PQI_ERR_HIP, "bf16 split failed");
    }
    if (g->use_adj) {
        std::vector<int32_t> ptr(n + 1, 0), idx;
        std::vector<double> wv;
        idx.reserve(nnz);
        wv.reserve(nnz);
        for (int64_t i = 0; i < n; ++i) {
            for (int64_t j = 0; j < n; ++j) {
                const double a = read_w(W, dtype, i * n + j);
                if (a != 0.0) {
                    idx.push_back((int32_t)j);
                    wv.push_back(a);
                }
            }
            ptr[i + 1] = (int32_t)idx.size();
        }
        int r = upload_adjacency(g, ptr, idx, wv);
        if (r != PQI_OK) {
            free_graph(g);
            return r;
        }
    }
    *out = g;
    return PQI_OK;
}

int pqi_graph_csr(pqi_ctx* c, const int64_t* indptr, const int64_t* indices, const void* data, int dtype,
                  int64_t n, int64_t nnz, pqi_graph** out) {
    if (!c || !indptr || !out || (nnz > 0 && (!indices || !data))) return fail(PQI_ERR_INVALID, "null argument");
    if (!valid_dtype(dtype)) return fail(PQI_ERR_INVALID, "unsupported dtype");
    if (n < 1 || n >= INT32_MAX) return fail(PQI_ERR_INVALID, "n out of range");
    if (indptr[0] != 0 || indptr[n] != nnz) return fail(PQI_ERR_INVALID, "indptr does not match nnz");
    *out = nullptr;
    CHK(set_device(c));
    // symmetrise: count degrees
    std::vector<int64_t> deg(n, 0);
    double maxabs = 0.0;
    bool all_int = true;
    long double w_up_l = 0.0L;
    for (int64_t u = 0; u < n; ++u) {
        if (indptr[u + 1] < indptr[u]) return fail(PQI_ERR_INVALID, "indptr is not monotone");
        for (int64_t e = indptr[u]; e < indptr[u + 1]; ++e) {
            int64_t v = indices[e];
            if (v <= u || v >= n) return fail(PQI_ERR_INVALID, "CSR must be strictly upper triangular");
            double a = read_w(data, dtype, e);
            if (std::isnan(a) || std::isinf(a)) return fail(PQI_ERR_INVALID, "non-finite weight");
            if (a == 0.0) continue;
            deg[u]++;
            deg[v]++;
            maxabs = std::max(maxabs, std::fabs(a));
            if (a != std::nearbyint(a)) all_int = false;
            w_up_l += a;
        }
    }
    std::vector<int32_t> ptr(n + 1, 0);
    for (int64_t u = 0; u < n; ++u) ptr[u + 1] = ptr[u] + (int32_t)deg[u];
    const int64_t tot = ptr[n];
    std::vector<int32_t> idx(tot);
    std::vector<double> wv(tot);
    std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
    for (int64_t u = 0; u < n; ++u)
        for (int64_t e = indptr[u]; e < indptr[u + 1]; ++e) {
            double a = read_w(data, dtype, e);
            if (a == 0.0) continue;
            int64_t v = indices[e];
            idx[fill[u]] = (int32_t)v;
            wv[fill[u]++] = a;
            idx[fill[v]] = (int32_t)u;
            wv[fill[v]++] = a;
        }
    // sort each row by column; reject duplicates
    for (int64_t u = 0; u < n; ++u) {
        std::vector<std::pair<int32_t, double>> rowv;
        for (int32_t e = ptr[u]; e < ptr[u + 1]; ++e) rowv.emplace_back(idx[e], wv[e]);
        std::sort(rowv.begin(), rowv.end(), [](auto& a, auto& b) { return a.first < b.first; });
        for (size_t t = 0; t < rowv.size(); ++t) {
            if (t && rowv[t].first == rowv[t - 1].first) return fail(PQI_ERR_INVALID, "duplicate CSR entry");
            idx[ptr[u] + t] = rowv[t].first;
            wv[ptr[u] + t] = rowv[t].second;
        }
    }
    pqi_graph* g = new pqi_graph();
    g->ctx = c;
    g->n = n;
    g->kind = classify(maxabs, all_int, dtype);
    g->n_pad = round_up(n, 128);
    g->w_up = (double)w_up_l;
    g->w_up_i = g->kind == PQI_WEIGHTS_I8 ? (int64_t)std::llround((double)w_up_l) : 0;
    g->use_adj = true;
    int r = upload_adjacency(g, ptr, idx, wv);
    if (r != PQI_OK) {
        free_graph(g);
        return r;
    }
    *out = g;
    return PQI_OK;
}

int pqi_graph_free(pqi_graph* g) {
    if (!g) return PQI_OK;
    (void)hipSetDevice(g->ctx->device);
    (void)hipStreamSynchronize(g->ctx->stream);
    ls_graphs_reset(g->ctx);  // captured launches may reference this graph
    free_graph(g);
    return PQI_OK;
}

int pqi_graph_get_info(const pqi_graph* g, pqi_graph_info* info) {
    if (!g || !info) return fail(PQI_ERR_INVALID, "null argument");
    info->n = g->n;
    info->weights = g->kind;
    info->has_dense = g->W != nullptr;
    info->nnz = g->nnz;
    info->w_up = g->w_up;
    info->device_bytes = g->device_bytes;
    return PQI_OK;
}

// ---------------------------------------------------------------------------
// State upload and scoring helpers
// ---------------------------------------------------------------------------

static int load_state(pqi_graph* g, const uint8_t* theta, uint8_t** x, void** h) {
    pqi_ctx* c = g->ctx;
    CHK(ensure(c, S_X, g->n, (void**)x));
    CHK(ensure(c, S_H, g->n * field_bytes(*g), h));
    for (int64_t i = 0; i < g->n; ++i)
        if (theta[i] > 1) return fail(PQI_ERR_INVALID, "theta entries must be 0 or 1");
    HIPCHK(hipMemcpyAsync(*x, theta, g->n, hipMemcpyHostToDevice, c->stream));
    launch_field(*g, *x, *h, c->stream);
    return PQI_OK;
}

// Pinned host staging of at least `bytes` (grown on demand; freed by pqi_destroy).
static int pin_ensure(pqi_ctx* c, int which, size_t bytes) {
    Pinned& pb = c->pin[which];
    if (pb.bytes >= bytes) return PQI_OK;
    if (pb.p) {
        HIPCHK(hipStreamSynchronize(which == 3 ? c->stream2 : c->stream));
        (void)hipHostFree(pb.p);
        pb = Pinned{};
    }
    const size_t want = std::max<size_t>(bytes, 4096);
    // coherent: the local search's reports are read while kernels still run
    HIPCHK(hipHostMalloc(&pb.p, want, hipHostMallocCoherent));
    std::memset(pb.p, 0, want);
    pb.bytes = want;
    HIPCHK(hipHostGetDevicePointer(&pb.dev, pb.p, 0));
    return PQI_OK;
}

// Waits until the local search report in pinned memory carries the number
// `want` (written last by the report kernel, after a system-scope fence).  The
// stream is queried now and then so that a failed or drained stream ends the
// wait with an error instead of a hang.
static int wait_report(pqi_ctx* c, const LsReport* r, unsigned long long want) {
    for (uint32_t spins = 1;; ++spins) {
        if (__atomic_load_n(&r->seq, __ATOMIC_ACQUIRE) == want) return PQI_OK;
        if ((spins & 0x3FFF) == 0) {
            const hipError_t e = hipStreamQuery(c->stream);
            if (e == hipSuccess) {
                if (__atomic_load_n(&r->seq, __ATOMIC_ACQUIRE) == want) return PQI_OK;
                return fail(PQI_ERR_HIP, "local search: an iteration report is missing");
            }
            if (e != hipErrorNotReady) return fail(PQI_ERR_HIP, std::string("local search: ") + hipGetErrorString(e));
            (void)hipGetLastError();  // "not ready" is no error of the calls that follow
        }
        __builtin_ia32_pause();
    }
}

// A report read whole: its number, and a check word that matches its record
// and its state (n bytes) as read now.
static bool report_whole(const LsReport* r, const uint8_t* xo, int64_t n, unsigned long long want) {
    LsReport loc;
    std::memcpy(&loc, r, sizeof(loc));
    if (loc.seq != want) return false;
    unsigned long long xs = 0;
    const int64_t n8 = n >> 3;
    for (int64_t j = 0; j < n8; ++j) {
        unsigned long long w;
        std::memcpy(&w, xo + 8 * j, 8);
        xs += ls_state_term(w, (unsigned long long)j);
    }
    if (n & 7) {
        unsigned long long t = 0;
        for (int64_t v = n8 << 3; v < n; ++v) t |= (unsigned long long)xo[v] << (8 * (v - (n8 << 3)));
        xs += ls_state_term(t, (unsigned long long)n8);
    }
    return ls_report_check(loc, xs, want) == loc.check;
}

// Waits until the report numbered `want` reads whole; after about a thousand
// tries the stream is drained and the report read once more (then a mismatch
// is an error).
static int settle_report(pqi_ctx* c, const LsReport* r, const uint8_t* xo, int64_t n, unsigned long long want) {
    for (uint32_t spins = 1;; ++spins) {
        if (report_whole(r, xo, n, want)) return PQI_OK;
        if ((spins & 0x3FF) == 0) {
            HIPCHK(hipStreamSynchronize(c->stream));
            if (report_whole(r, xo, n, want)) return PQI_OK;
            return fail(PQI_ERR_HIP, "local search: an iteration report does not match its check word");
        }
        __builtin_ia32_pause();
    }
}

static int score_device(pqi_graph* g, const uint8_t* x, double* cut, double* en) {
    pqi_ctx* c = g->ctx;
    double* out;
    void* hb;
    CHK(ensure(c, S_SCORE, 16, (void**)&out));
    CHK(ensure(c, S_HB, g->n * field_bytes(*g), &hb));
    launch_score_with(*g, x, 1, out, out + 1, hb, c->stream);
    double hv[2];
    HIPCHK(hipMemcpyAsync(hv, out, 16, hipMemcpyDeviceToHost, c->stream));
    CHK(sync(c));
    *cut = hv[0];
    *en = hv[1];
    return PQI_OK;
}

static int read_out(pqi_ctx* c, const PhaseOut* d, PhaseOut* hst) {
    HIPCHK(hipMemcpyAsync(hst, d, sizeof(PhaseOut), hipMemcpyDeviceToHost, c->stream));
    return sync(c);
}

int pqi_score(pqi_graph* g, const uint8_t* theta, double* cut, double* energy) {
    if (!g || !theta || !cut || !energy) return fail(PQI_ERR_INVALID, "null argument");
    CHK(set_device(g->ctx));
    uint8_t* x;
    void* h;
    CHK(load_state(g, theta, &x, &h));
    return score_device(g, x, cut, energy);
}

static int eval_on_device(pqi_graph* g, uint32_t* dbits, int64_t B, int is_spin_glass, bool bf16, double* values,
                          int64_t* best_index, double* best_value, double* max_rel_err);

int pqi_eval_batch_ex(pqi_graph* g, const uint32_t* bits, int64_t B, int is_spin_glass, int engine,
                      double* values, int64_t* best_index, double* best_value, double* max_rel_err) {
    if (!g || (!bits && B > 0)) return fail(PQI_ERR_INVALID, "null argument");
    if (B < 1) return fail(PQI_ERR_INVALID, "B must be >= 1");
    if (engine != PQI_ENGINE_AUTO && engine != PQI_ENGINE_EXACT && engine != PQI_ENGINE_BF16)
        return fail(PQI_ERR_INVALID, "bad engine");
    const bool bf16 = g->Wbf != nullptr && engine != PQI_ENGINE_EXACT;
    if (engine == PQI_ENGINE_BF16 && !bf16)
        return fail(PQI_ERR_UNSUPPORTED, "the split-BF16 engine needs dense fp32 weights");
    pqi_ctx* c = g->ctx;
    CHK(set_device(c));
    const int64_t nw = (g->n + 31) >> 5;
    const int64_t wpr = g->n_pad >> 5;
    uint32_t* dbits;
    CHK(ensure(c, S_BITS, (size_t)B * wpr * 4, (void**)&dbits));
    // padding words are zero: padded columns of W are zero, so they add nothing
    CHK(upload_rows(c, dbits, (size_t)wpr * 4, bits, (size_t)nw * 4, (size_t)B));
    return eval_on_device(g, dbits, B, is_spin_glass, bf16, values, best_index, best_value, max_rel_err);
}

// Scores B candidate rows already on the device (n_pad/32 words per row).
static int eval_on_device(pqi_graph* g, uint32_t* dbits, int64_t B, int is_spin_glass, bool bf16, double* values,
                          int64_t* best_index, double* best_value, double* max_rel_err) {
    pqi_ctx* c = g->ctx;
    const int64_t wpr = g->n_pad >> 5;
    double* q;
    void* scratch;
    CHK(ensure(c, S_Q, (size_t)B * 8, (void**)&q));
    CHK(ensure(c, S_BATCH, bf16 ? batch_bf16_scratch_bytes(*g, B) : batch_scratch_bytes(*g, B), &scratch));
    if (bf16) launch_batch_bf16(*g, dbits, wpr, B, q, scratch, c->stream);
    else launch_batch_quadratic(*g, dbits, wpr, B, q, scratch, c->stream);
    std::vector<double> hq(B);
    HIPCHK(hipMemcpyAsync(hq.data(), q, B * 8, hipMemcpyDeviceToHost, c->stream));
    CHK(sync(c));
    // cut = (2 W_up - q)/4, E = 2 cut - W_up = -q/2
    auto value_of = [&](double qv) {
        if (g->kind == PQI_WEIGHTS_I8) {
            const long long qi = (long long)qv;
            return is_spin_glass ? (double)(-qi / 2) : (double)((2 * g->w_up_i - qi) / 4);
        }
        return is_spin_glass ? -qv / 2.0 : (2.0 * g->w_up - qv) / 4.0;
    };
    std::vector<double> v(B);
    for (int64_t b = 0; b < B; ++b) v[b] = value_of(hq[b]);
    int64_t best = 0;
    if (bf16) {
        // FP64 re-score of the K best approximate values (one pass over W)
        const int64_t K = std::min<int64_t>(16, B);
        std::vector<int64_t> order(B);
        for (int64_t b = 0; b < B; ++b) order[b] = b;
        std::partial_sort(order.begin(), order.begin() + K, order.end(), [&](int64_t a, int64_t b2) {
            return v[a] > v[b2] || (v[a] == v[b2] && a < b2);
        });
        order.resize(K);
        std::sort(order.begin(), order.end());
        int64_t* dsel;
        uint32_t* rows;
        double* q64;
        void* sc2;
        CHK(ensure(c, S_SEL, K * 8, (void**)&dsel));
        CHK(ensure(c, S_RANK, (size_t)K * wpr * 4, (void**)&rows));
        CHK(ensure(c, S_QB, K * 8, (void**)&q64));
        CHK(ensure(c, S_DBL, std::max(batch_scratch_bytes(*g, K), double_scratch_bytes(*g)), &sc2));
        HIPCHK(hipMemcpyAsync(dsel, order.data(), K * 8, hipMemcpyHostToDevice, c->stream));
        launch_gather_rows(dbits, wpr, dsel, K, rows, c->stream);
        launch_batch_quadratic(*g, rows, wpr, K, q64, sc2, c->stream);
        std::vector<double> hq64(K);
        HIPCHK(hipMemcpyAsync(hq64.data(), q64, K * 8, hipMemcpyDeviceToHost, c->stream));
        CHK(sync(c));
        double err = 0.0, bv = -INFINITY;
        for (int64_t t = 0; t < K; ++t) {
            const double exact = value_of(hq64[t]);
            const double den = std::max(std::fabs(exact), 1e-300);
            err = std::max(err, std::fabs(v[order[t]] - exact) / den);
            v[order[t]] = exact;
            if (exact > bv) { bv = exact; best = order[t]; }
        }
        if (max_rel_err) *max_rel_err = err;
    } else {
        for (int64_t b = 1; b < B; ++b)
            if (v[b] > v[best]) best = b;
        if (max_rel_err) *max_rel_err = 0.0;
    }
    if (values) std::memcpy(values, v.data(), B * 8);
    if (best_index) *best_index = best;
    if (best_value) *best_value = v[best];
    return PQI_OK;
}

int pqi_eval_batch(pqi_graph* g, const uint32_t* bits, int64_t B, int is_spin_glass, double* values,
                   int64_t* best_index, double* best_value) {
    return pqi_eval_batch_ex(g, bits, B, is_spin_glass, PQI_ENGINE_AUTO, values, best_index, best_value, nullptr);
}

static int sample_to_device(pqi_graph* g, const double* cum_prob, int64_t ncum, double repulsion_base, uint64_t seed,
                            int64_t shot_offset, int64_t shots, uint32_t** dbits) {
    if (!cum_prob || ncum < 1) return fail(PQI_ERR_INVALID, "cum_prob must have at least one entry");
    if (shots < 1 || shot_offset < 0) return fail(PQI_ERR_INVALID, "shots must be >= 1, shot_offset >= 0");
    if (!(repulsion_base > 0.0) || !std::isfinite(repulsion_base))
        return fail(PQI_ERR_INVALID, "repulsion_base must be positive and finite");
    pqi_ctx* c = g->ctx;
    const int64_t wpr = g->n_pad >> 5;
    double* dcum;
    CHK(ensure(c, S_CUM, (size_t)ncum * 8, (void**)&dcum));
    CHK(ensure(c, S_BITS, (size_t)shots * wpr * 4, (void**)dbits));
    HIPCHK(hipMemcpyAsync(dcum, cum_prob, (size_t)ncum * 8, hipMemcpyHostToDevice, c->stream));
    return launch_sampler(*g, dcum, ncum, repulsion_base, seed, shot_offset, shots, *dbits, wpr, c->stream);
}

int pqi_sample_candidates(pqi_graph* g, const double* cum_prob, int64_t ncum, double repulsion_base, uint64_t seed,
                          int64_t shot_offset, int64_t shots, uint8_t* out) {
    if (!g || !out) return fail(PQI_ERR_INVALID, "null argument");
    pqi_ctx* c = g->ctx;
    CHK(set_device(c));
    uint32_t* dbits;
    CHK(sample_to_device(g, cum_prob, ncum, repulsion_base, seed, shot_offset, shots, &dbits));
    const int64_t wpr = g->n_pad >> 5, n = g->n;
    std::vector<uint32_t> hb((size_t)shots * wpr);
    HIPCHK(hipMemcpyAsync(hb.data(), dbits, hb.size() * 4, hipMemcpyDeviceToHost, c->stream));
    CHK(sync(c));
    for (int64_t s2 = 0; s2 < shots; ++s2)
        for (int64_t v = 0; v < n; ++v) out[s2 * n + v] = (uint8_t)((hb[s2 * wpr + (v >> 5)] >> (v & 31)) & 1u);
    return PQI_OK;
}

int pqi_sample_round(pqi_graph* g, const double* cum_prob, int64_t ncum, double repulsion_base, uint64_t seed,
                     int64_t shot_offset, int64_t shots, int is_spin_glass, int64_t* best_index, double* best_value,
                     uint8_t* best_bits) {
    if (!g || !best_index || !best_value || !best_bits) return fail(PQI_ERR_INVALID, "null argument");
    pqi_ctx* c = g->ctx;
    CHK(set_device(c));
    uint32_t* dbits;
    CHK(sample_to_device(g, cum_prob, ncum, repulsion_base, seed, shot_offset, shots, &dbits));
    CHK(eval_on_device(g, dbits, shots, is_spin_glass, false, nullptr, best_index, best_value, nullptr));
    const int64_t wpr = g->n_pad >> 5, n = g->n;
    std::vector<uint32_t> row(wpr);
    HIPCHK(hipMemcpyAsync(row.data(), dbits + *best_index * wpr, wpr * 4, hipMemcpyDeviceToHost, c->stream));
    CHK(sync(c));
    for (int64_t v = 0; v < n; ++v) best_bits[v] = (uint8_t)((row[v >> 5] >> (v & 31)) & 1u);
    return PQI_OK;
}

int pqi_single_flip_best(pqi_graph* g, const uint8_t* theta, int is_spin_glass, int64_t* k, double* gain) {
    if (!g || !theta || !k || !gain) return fail(PQI_ERR_INVALID, "null argument");
    pqi_ctx* c = g->ctx;
    CHK(set_device(c));
    uint8_t* x;
    void* h;
    CHK(load_state(g, theta, &x, &h));
    PhaseOut* d;
    CHK(ensure(c, S_OUT, sizeof(PhaseOut), (void**)&d));
    launch_single_best(*g, x, h, d, c->stream);
    PhaseOut o;
    CHK(read_out(c, d, &o));
    *k = o.k;
    *gain = is_spin_glass ? 2.0 * o.gain : o.gain;
    return PQI_OK;
}

int pqi_single_flip_descent(pqi_graph* g, uint8_t* theta, int is_spin_glass, int64_t max_steps, int64_t* accepts,
                            double* gain) {
    if (!g || !theta || !accepts || !gain) return fail(PQI_ERR_INVALID, "null argument");
    if (max_steps < 0) return fail(PQI_ERR_INVALID, "max_steps must be >= 0");
    pqi_ctx* c = g->ctx;
    CHK(set_device(c));
    uint8_t* x;
    void* h;
    CHK(load_state(g, theta, &x, &h));
    PhaseOut* d;
    CHK(ensure(c, S_OUT, sizeof(PhaseOut), (void**)&d));
    // bounded launches keep every kernel short
    const int64_t cap = 16 * g->n + 4096;
    int64_t done = 0;
    double sum = 0.0;
    PhaseOut o;
    do {
        const int64_t want = std::min(cap, max_steps - done);
        launch_descent(*g, x, h, want, d, c->stream);
        CHK(read_out(c, d, &o));
        done += o.accepts;
        sum += o.gain_sum;
        if (o.accepts < want || want == 0) break;
    } while (done < max_steps);
    HIPCHK(hipMemcpyAsync(theta, x, g->n, hipMemcpyDeviceToHost, c->stream));
    CHK(sync(c));
    *accepts = done;
    *gain = is_spin_glass ? 2.0 * sum : sum;
    return PQI_OK;
}

int pqi_double_flip_best(pqi_graph* g, const uint8_t* theta, int is_spin_glass, int64_t thread_count, int mode,
                         int64_t* k, int64_t* l, double* gain) {
    if (!g || !theta || !k || !l || !gain) return fail(PQI_ERR_INVALID, "null argument");
    if (mode != PQI_DOUBLE_REFERENCE && mode != PQI_DOUBLE_EXHAUSTIVE) return fail(PQI_ERR_INVALID, "bad mode");
    if (thread_count < 1) return fail(PQI_ERR_INVALID, "thread_count must be >= 1");
    if (g->n < 2) return fail(PQI_ERR_INVALID, "n must be >= 2");
    pqi_ctx* c = g->ctx;
    CHK(set_device(c));
    uint8_t* x;
    void* h;
    CHK(load_state(g, theta, &x, &h));
    PhaseOut* d;
    void* ds;
    CHK(ensure(c, S_OUT, sizeof(PhaseOut), (void**)&d));
    CHK(ensure(c, S_DBL, double_scratch_bytes(*g), &ds));
    launch_double_best(*g, x, h, thread_count, mode, d, ds, c->stream);
    PhaseOut o;
    CHK(read_out(c, d, &o));
    *k = o.k;
    *l = o.l;
    *gain = is_spin_glass ? 2.0 * o.gain : o.gain;
    return PQI_OK;
}

}  // extern "C"

namespace pqi {
int ensure_pinned(pqi_ctx* c, int which, size_t bytes) { return pin_ensure(c, which, bytes); }
}  // namespace pqi

namespace pqi {

struct GrayBufs {
    uint8_t* wx;
    void* wh;
    double* we;
    int64_t* wev;
    int64_t* sel;
    double* seld;
    double* delta;
};

static int check_params(const pqi_graph* g, const pqi_search_params* p) {
    if (!p) return fail(PQI_ERR_INVALID, "params is null");
    if (p->walkers < 1) return fail(PQI_ERR_INVALID, "walkers must be >= 1");
    if (p->seed_multiple < 1) return fail(PQI_ERR_INVALID, "seed_multiple must be >= 1");
    if (p->gray_iterations < 0) return fail(PQI_ERR_INVALID, "gray_iterations must be >= 0");
    if (p->double_mode != PQI_DOUBLE_REFERENCE && p->double_mode != PQI_DOUBLE_EXHAUSTIVE)
        return fail(PQI_ERR_INVALID, "bad double_mode");
    int64_t nb = (g->n + 63) >> 6;
    if (p->walkers > p->walkers * p->seed_multiple * nb) return fail(PQI_ERR_INVALID, "more walkers than seeds");
    if (p->walkers > (1 << 20)) return fail(PQI_ERR_INVALID, "walkers too large");
    return PQI_OK;
}

// Score every Gray seed of (x, h) and select the top `walkers` (stable order).
// Allocate every buffer gray_seeds uses (so that no allocation happens while
// kernels are queued).
static int gray_seeds_buffers(pqi_graph* g, const pqi_search_params* p, GrayBufs& b) {
    pqi_ctx* c = g->ctx;
    const int64_t nb = (g->n + 63) >> 6;
    const int64_t count = p->walkers * p->seed_multiple * nb;
    CHK(ensure(c, S_DELTA, count * 8, (void**)&b.delta));
    CHK(ensure(c, S_SEL, p->walkers * 8, (void**)&b.sel));
    CHK(ensure(c, S_SELD, p->walkers * 8, (void**)&b.seld));
    if (g->kind == PQI_WEIGHTS_I8 && g->W != nullptr && p->seed_engine != PQI_SEEDS_LOWRANK) {
        const int64_t wpr = g->n_pad >> 5;
        void* t;
        CHK(ensure(c, S_BITS, (size_t)count * wpr * 4, &t));
        CHK(ensure(c, S_BATCH, batch_scratch_bytes(*g, count + 1), &t));
    }
    void* rank;
    CHK(ensure(c, S_RANK, topk_scratch_bytes(count), &rank));
    return PQI_OK;
}

// Gray seeds of the state: their gains and the top `walkers` (sel, seld).  With
// a gate, the seed rows and the int8 MFMA contraction do nothing while *gate is
// nonzero (their results are then not used).  With qi_out, the int8 MFMA engine
// selects straight from the row totals, fills sel only and returns the totals
// (the caller computes seld; nullptr when that engine does not apply).
static int gray_seeds(pqi_graph* g, const uint8_t* x, const void* h, const pqi_search_params* p, GrayBufs& b,
                      int64_t* seed_count, const int32_t* gate = nullptr, unsigned long long** qi_out = nullptr) {
    if (qi_out) *qi_out = nullptr;
    pqi_ctx* c = g->ctx;
    const int64_t nb = (g->n + 63) >> 6;
    const int64_t bs = p->walkers * p->seed_multiple;
    const int64_t count = bs * nb;
    *seed_count = count;
    CHK(ensure(c, S_DELTA, count * 8, (void**)&b.delta));
    CHK(ensure(c, S_SEL, p->walkers * 8, (void**)&b.sel));
    CHK(ensure(c, S_SELD, p->walkers * 8, (void**)&b.seld));
    bool mfma = g->kind == PQI_WEIGHTS_I8 && g->W != nullptr &&
                (p->seed_engine == PQI_SEEDS_MFMA || p->seed_engine == PQI_SEEDS_AUTO);
    if (p->seed_engine == PQI_SEEDS_MFMA && !mfma)
        return fail(PQI_ERR_UNSUPPORTED, "MFMA seed scoring needs dense int8 weights");
    if (mfma) {
        const int64_t wpr = g->n_pad >> 5;
        uint32_t* bits;
        void* scratch;
        CHK(ensure(c, S_BITS, (size_t)count * wpr * 4, (void**)&bits));
        CHK(ensure(c, S_BATCH, batch_scratch_bytes(*g, count + 1), &scratch));
        void* rank;
        CHK(ensure(c, S_RANK, topk_scratch_bytes(count), &rank));
        unsigned long long* qi = (unsigned long long*)scratch;
        if (qi_out) {
            // seed rows (clearing the row totals and ranks), int8 MFMA
            // contraction, top-k by row total
            launch_seed_bits(x, g->n, wpr, bs, count, bits, c->stream, qi, gate, (int32_t*)rank);
            launch_batch_i8_raw(*g, bits, wpr, count, qi, c->stream, gate);
            launch_topk_q(qi, count, p->walkers, (int32_t*)rank, b.sel, c->stream);
            *qi_out = qi;
            return PQI_OK;
        }
        // seed rows (clearing the row totals), int8 MFMA contraction, cut
        // differences against q_base = s^T h (clearing the ranks), top-k
        launch_seed_bits(x, g->n, wpr, bs, count, bits, c->stream, qi, gate);
        launch_batch_i8_raw(*g, bits, wpr, count, qi, c->stream, gate);
        launch_seed_delta_i8(x, (const int32_t*)h, g->n, qi, count, b.delta, (int32_t*)rank, c->stream);
        launch_topk_gather(b.delta, count, p->walkers, b.sel, b.seld, (int32_t*)rank, c->stream);
        return PQI_OK;
    }
    launch_seed_lowrank(*g, x, h, bs, count, b.delta, c->stream);
    void* rank;
    CHK(ensure(c, S_RANK, topk_scratch_bytes(count), &rank));
    launch_topk(b.delta, count, p->walkers, b.sel, rank, c->stream);
    launch_gather(b.delta, b.sel, p->walkers, b.seld, c->stream);
    return PQI_OK;
}

// The device loop's fused Gray seeds (reference-mode double flips, int8 MFMA
// seeds): the seed rows are written by k_ls_descent_double's extra work groups,
// then the MFMA contraction and the ranking run, and k_ls_seed_take takes the
// top G.  Fills the rows and returns true when this applies; no launches.
static bool fused_seed_plan(pqi_graph* g, const pqi_search_params* p, SeedRows* sr, unsigned long long** qi,
                            int32_t** rank) {
    pqi_ctx* c = g->ctx;
    const bool mfma = g->kind == PQI_WEIGHTS_I8 && g->W != nullptr &&
                      (p->seed_engine == PQI_SEEDS_MFMA || p->seed_engine == PQI_SEEDS_AUTO);
    if (!mfma || p->double_mode != PQI_DOUBLE_REFERENCE) return false;
    const int64_t nb = (g->n + 63) >> 6;
    sr->bs = p->walkers * p->seed_multiple;
    sr->count = sr->bs * nb;
    sr->wpr = g->n_pad >> 5;
    void *bits, *scratch, *rk;
    // the buffers exist already (gray_seeds_buffers): these calls only look them up
    if (ensure(c, S_BITS, (size_t)sr->count * sr->wpr * 4, &bits) != PQI_OK ||
        ensure(c, S_BATCH, batch_scratch_bytes(*g, sr->count + 1), &scratch) != PQI_OK ||
        ensure(c, S_RANK, topk_scratch_bytes(sr->count), &rk) != PQI_OK)
        return false;
    sr->bits = (uint32_t*)bits;
    sr->qzero = (unsigned long long*)scratch;
    sr->rzero = (int32_t*)rk;
    *qi = sr->qzero;
    *rank = sr->rzero;
    return true;
}

static int gray_walk(pqi_graph* g, const uint8_t* x, const void* h, const pqi_search_params* p, GrayBufs& b,
                     int64_t walkers, int64_t T) {
    pqi_ctx* c = g->ctx;
    const size_t fb = field_bytes(*g);
    CHK(ensure(c, S_WX, (size_t)walkers * g->n, (void**)&b.wx));
    CHK(ensure(c, S_WH, (size_t)walkers * g->n * fb, &b.wh));
    CHK(ensure(c, S_WE, walkers * 8, (void**)&b.we));
    CHK(ensure(c, S_WEV, walkers * 8, (void**)&b.wev));
    launch_gray_walk(*g, x, h, b.sel, b.seld, p->walkers * p->seed_multiple, walkers, T, b.wx, b.wh, b.we, b.wev,
                     c->stream);
    return PQI_OK;
}

// Optional event timing of one kernel group on the context stream.  The
// events only time: they are created without the system-scope fence (no cache
// writeback and invalidation when they are recorded); the host reads them
// after it has seen later work of the stream complete.
struct EventTimer {
    bool on = false;
    hipEvent_t a = nullptr, b = nullptr;
    explicit EventTimer(bool enable) : on(enable) {
        if (on && (hipEventCreateWithFlags(&a, hipEventDisableSystemFence) != hipSuccess ||
                   hipEventCreateWithFlags(&b, hipEventDisableSystemFence) != hipSuccess))
            on = false;
    }
    ~EventTimer() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
    void start(hipStream_t s) {
        if (on) (void)hipEventRecord(a, s);
    }
    void stop(hipStream_t s) {
        if (on) (void)hipEventRecord(b, s);
    }
    // events for a kernel launch that records them itself (hipExtLaunchKernel)
    hipEvent_t start_event() const { return on ? a : nullptr; }
    hipEvent_t stop_event() const { return on ? b : nullptr; }
    // call once later work of the stream is known to be complete (the stop
    // event is then recorded; waiting on it returns at once)
    double seconds() {
        float ms = 0.f;
        if (!on) return 0.0;
        if (hipEventSynchronize(b) == hipSuccess && hipEventElapsedTime(&ms, a, b) == hipSuccess) return ms * 1e-3;
        (void)hipGetLastError();
        return 0.0;
    }
};

// Gray steps one walker of the reference performs in T iterations: iteration
// i flips ctz(i+1) + 64 b for every block b, skipping positions >= n.
static int64_t reference_gray_steps(int64_t T, int64_t n) {
    const int64_t nb = (n + 63) >> 6;
    int64_t total = 0;
    for (int p = 0; p < 63; ++p) {
        const int64_t cnt = (T >> p) - (T >> (p + 1));
        if (!cnt) continue;
        int64_t valid = 0;
        for (int64_t b = 0; b < nb; ++b) valid += (p + 64 * b < n) ? 1 : 0;
        total += cnt * valid;
    }
    return total;
}

static double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace pqi

extern "C" {

int pqi_gray_search(pqi_graph* g, const uint8_t* theta, const pqi_search_params* p, uint8_t* state_out,
                    double* seed_gain, double* walk_gain) {
    if (!g || !theta || !state_out || !seed_gain || !walk_gain) return fail(PQI_ERR_INVALID, "null argument");
    CHK(check_params(g, p));
    pqi_ctx* c = g->ctx;
    CHK(set_device(c));
    uint8_t* x;
    void* h;
    CHK(load_state(g, theta, &x, &h));
    GrayBufs b;
    int64_t count;
    CHK(gray_seeds(g, x, h, p, b, &count));
    const double scale = p->is_spin_glass ? 2.0 : 1.0;
    double top;
    HIPCHK(hipMemcpyAsync(&top, b.seld, 8, hipMemcpyDeviceToHost, c->stream));
    CHK(sync(c));
    *seed_gain = scale * top;
    *walk_gain = 0.0;
    if (top > 0.0) {
        CHK(gray_walk(g, x, h, p, b, 1, 0));
        HIPCHK(hipMemcpyAsync(state_out, b.wx, g->n, hipMemcpyDeviceToHost, c->stream));
        return sync(c);
    }
    const int64_t T = (p->gray_iterations + p->walkers - 1) / p->walkers;
    CHK(gray_walk(g, x, h, p, b, p->walkers, T));
    std::vector<double> we(p->walkers);
    HIPCHK(hipMemcpyAsync(we.data(), b.we, p->walkers * 8, hipMemcpyDeviceToHost, c->stream));
    CHK(sync(c));
    int64_t bw = 0;
    for (int64_t w = 1; w < p->walkers; ++w)
        if (we[w] > we[bw]) bw = w;
    *walk_gain = scale * we[bw];
    HIPCHK(hipMemcpyAsync(state_out, b.wx + bw * g->n, g->n, hipMemcpyDeviceToHost, c->stream));
    return sync(c);
}

static void field_repair_stats(const FieldRepair& f, pqi_search_stats* st) {
    st->field_checks = (int64_t)f.checks;
    st->field_repairs = (int64_t)f.entries;
    double d;
    std::memcpy(&d, &f.max_bits, sizeof(d));
    st->max_field_drift = d;
}

static int read_field_repair(pqi_ctx* c, const FieldRepair* fix, pqi_search_stats* st) {
    FieldRepair f;
    HIPCHK(hipMemcpyAsync(&f, fix, sizeof(f), hipMemcpyDeviceToHost, c->stream));
    CHK(sync(c));
    st->field_checks = (int64_t)f.checks;
    st->field_repairs = (int64_t)f.entries;
    double d;
    std::memcpy(&d, &f.max_bits, sizeof(d));
    st->max_field_drift = d;
    return PQI_OK;
}

// Counters and phase clocks of a device-decided search from its control record.
static void fill_search_stats(const pqi_ctx* c, const LsCtl& hc, pqi_search_stats* st) {
    int rate_khz = 0;
    if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, c->device) != hipSuccess || rate_khz <= 0)
        rate_khz = 100000;
    const double tick = 1e-3 / rate_khz;
    st->rounds = hc.rounds;
    st->max_energy = hc.max_energy;
    st->single_evals = hc.single_evals;
    st->single_accepts = hc.single_accepts;
    st->double_evals = hc.double_evals;
    st->double_accepts = hc.double_accepts;
    st->seed_evals = hc.seed_evals;
    st->seed_accepts = hc.seed_accepts;
    st->walk_accepts = hc.walk_accepts;
    st->reheat_accepts = hc.reheat_accepts;
    st->gray_evals = hc.gray_evals;
    st->gray_evals_computed = hc.gray_evals_computed;
    st->descent_evals = hc.descent_evals;
    st->seconds_single = hc.ticks[0] * tick;
    st->seconds_double = hc.ticks[1] * tick;
    st->seconds_seeds = hc.ticks[2] * tick;
    st->seconds_walk = hc.ticks[3] * tick;
    st->seconds_reheat = hc.ticks[4] * tick;
}

// Host-driven loop: one synchronisation per phase decision (PQI_FLAG_HOST_LOOP).
static int local_search_host(pqi_graph* g, uint8_t* theta, const pqi_search_params* p, pqi_search_stats* st);

// Device-driven loop: one iteration of the reference loop is enqueued as a
// whole and decided by control kernels (LsCtl); one synchronisation per
// iteration.  Same moves, counters and max_energy as the host-driven loop.
static int local_search_device(pqi_graph* g, uint8_t* theta, const pqi_search_params* p, pqi_search_stats* st) {
    pqi_ctx* c = g->ctx;
    const int64_t n = g->n;
    const size_t fb = field_bytes(*g);
    const int64_t C = n * (n - 1) / 2;
    const int64_t G = p->walkers;
    const int64_t T = (p->gray_iterations + G - 1) / G;
    const int64_t dbl_evals = p->double_mode == PQI_DOUBLE_EXHAUSTIVE ? C : std::min(G, C);

    static const bool iter_graph_enabled = [] {
        const char* e = getenv("PQI_LS_ITER_GRAPH");
        return !(e && e[0] == '0');
    }();
    // int8 weights with adjacency lists: the start is one launch reading the
    // caller's state from pinned memory, and every iteration ends with a report
    // (control record, repair counters, exact score, state) written into pinned
    // memory, so the search needs no copy before the first descent and nothing
    // after the last synchronisation
    const bool fast = iter_graph_enabled && g->kind == PQI_WEIGHTS_I8 && g->adj_ptr &&
                      (g->use_adj || g->W == nullptr) && n <= 65536 && !(p->flags & PQI_FLAG_TEST_DRIFT);
    // report buffer layout: two slots of (report, state)
    const size_t slot_bytes = ((sizeof(LsReport) + 255) & ~(size_t)255) + (((size_t)n + 255) & ~(size_t)255);
    const size_t off_rep[2] = {0, slot_bytes};
    const size_t off_xout[2] = {off_rep[0] + ((sizeof(LsReport) + 255) & ~(size_t)255),
                                off_rep[1] + ((sizeof(LsReport) + 255) & ~(size_t)255)};
    uint8_t *x, *rx;
    void *h, *rh;
    if (fast) {
        for (int64_t i = 0; i < n; ++i)
            if (theta[i] > 1) return fail(PQI_ERR_INVALID, "theta entries must be 0 or 1");
        CHK(pin_ensure(c, 0, (size_t)n));
        CHK(pin_ensure(c, 1, off_xout[1] + (size_t)n));
        if (!c->ls.seq_dev) {
            HIPCHK(hipMalloc((void**)&c->ls.seq_dev, sizeof(unsigned long long)));
            HIPCHK(hipMemsetAsync(c->ls.seq_dev, 0, sizeof(unsigned long long), c->stream));
            c->ls.seq_host = 0;
            c->ls.seq_dirty = false;
        } else if (c->ls.seq_dirty) {
            // a previous call ended early: take the count from the device
            HIPCHK(hipStreamSynchronize(c->stream));
            HIPCHK(hipMemcpy(&c->ls.seq_host, c->ls.seq_dev, sizeof(unsigned long long), hipMemcpyDeviceToHost));
            c->ls.seq_dirty = false;
        }
        std::memcpy(c->pin[0].p, theta, (size_t)n);
        CHK(ensure(c, S_X, n, (void**)&x));
        CHK(ensure(c, S_H, n * fb, &h));
    } else {
        CHK(load_state(g, theta, &x, &h));
    }
    LsReport* rep[2] = {nullptr, nullptr};
    LsReport* rep_dev[2] = {nullptr, nullptr};
    uint8_t* xout[2] = {nullptr, nullptr};
    uint8_t* xout_dev[2] = {nullptr, nullptr};
    if (fast)
        for (int k = 0; k < 2; ++k) {
            rep[k] = (LsReport*)((char*)c->pin[1].p + off_rep[k]);
            rep_dev[k] = (LsReport*)((char*)c->pin[1].dev + off_rep[k]);
            xout[k] = (uint8_t*)c->pin[1].p + off_xout[k];
            xout_dev[k] = (uint8_t*)c->pin[1].dev + off_xout[k];
        }
    PhaseOut *dout, *dout2;
    void* dsc;
    LsCtl* ctl;
    FieldRepair* fix;
    CHK(ensure(c, S_FIX, sizeof(FieldRepair), (void**)&fix));
    if (!fast) HIPCHK(hipMemsetAsync(fix, 0, sizeof(FieldRepair), c->stream));
    if (p->flags & PQI_FLAG_TEST_DRIFT) launch_perturb_field(*g, h, c->stream);
    CHK(ensure(c, S_RX, n, (void**)&rx));
    CHK(ensure(c, S_RH, n * fb, &rh));
    CHK(ensure(c, S_OUT, sizeof(PhaseOut), (void**)&dout));
    CHK(ensure(c, S_OUT2, sizeof(PhaseOut), (void**)&dout2));
    CHK(ensure(c, S_DBL, double_scratch_bytes(*g), &dsc));
    CHK(ensure(c, S_CTL, sizeof(LsCtl), (void**)&ctl));
    // every buffer of the iteration exists before the first launch (no
    // reallocation while kernels are queued)
    GrayBufs b;
    int64_t count;
    CHK(ensure(c, S_WX, (size_t)G * n, (void**)&b.wx));
    CHK(ensure(c, S_WH, (size_t)G * n * fb, &b.wh));
    CHK(ensure(c, S_WE, G * 8, (void**)&b.we));
    CHK(ensure(c, S_WEV, G * 8, (void**)&b.wev));
    CHK(gray_seeds_buffers(g, p, b));

    // the starting max_energy is the exact score of the start, set on the
    // device (no synchronisation before the first iteration)
    double* score;
    void* hb;
    CHK(ensure(c, S_SCORE, 16, (void**)&score));
    CHK(ensure(c, S_HB, n * fb, &hb));
    LsCtl hc;
    std::memset(&hc, 0, sizeof(hc));
    hc.scale = p->is_spin_glass ? 2.0 : 1.0;
    hc.phase = -1;
    hc.max_rounds = p->max_rounds;
    if (!fast) {
        HIPCHK(hipMemcpyAsync(ctl, &hc, sizeof(hc), hipMemcpyHostToDevice, c->stream));
        launch_score_with(*g, x, 1, score, score + 1, hb, c->stream);
        launch_ls_start_energy(ctl, score, p->is_spin_glass, c->stream);
        // the reheat state is read by speculative launches before its first use
        HIPCHK(hipMemcpyAsync(rx, x, n, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(rh, h, n * fb, hipMemcpyDeviceToDevice, c->stream));
    }

    const bool timing = (p->flags & PQI_FLAG_TIME_KERNELS) != 0;
    EventTimer walk_timer(timing), descent_timer(timing);
    const int64_t cap = 16 * n + 4096;  // bounded descent launches
    const int64_t gray_evals = G * reference_gray_steps(T, n);
    const int64_t bs = p->walkers * p->seed_multiple;
    // segment 0: after the descent, up to the main walk; segment 1: after it
    // Control steps that sit between single-work-group phases run fused in one
    // kernel each (a kernel boundary costs a few microseconds); the exhaustive
    // double flip is multi-work-group and keeps its own launches.
    const bool fused = p->double_mode == PQI_DOUBLE_REFERENCE;
    // int8 MFMA seeds with the fused control kernels: the seed rows are written
    // alongside the double flip, and the top G are taken by the seed decision
    SeedRows sr;
    unsigned long long* fqi = nullptr;
    int32_t* frank = nullptr;
    const bool fuse_seeds = fused && fused_seed_plan(g, p, &sr, &fqi, &frank);
    auto segment = [&](int which, const LsReportTo* rt) -> int {
        if (which == 0) {
            if (fused) {
                launch_ls_descent_double(*g, ctl, dout, cap, x, h, G, dout2, dbl_evals, c->stream,
                                         fuse_seeds ? &sr : nullptr);
            } else {
                launch_ls_after_descent(ctl, dout, cap, c->stream);
                launch_double_best(*g, x, h, G, p->double_mode, dout2, dsc, c->stream);
                launch_ls_after_double(*g, ctl, dout2, x, h, dbl_evals, c->stream);
            }
            unsigned long long* qi = nullptr;
            if (fuse_seeds) {
                launch_batch_i8_raw(*g, sr.bits, sr.wpr, sr.count, fqi, c->stream, &ctl->restart);
                launch_rank_q(fqi, sr.count, frank, c->stream);
                qi = fqi;
                count = sr.count;
            } else {
                CHK(gray_seeds(g, x, h, p, b, &count, &ctl->restart, &qi));
            }
            // the best seed, if it improves, is applied to the state
            launch_ls_seed_take(*g, ctl, b.seld, count, b.sel, bs, x, h, c->stream, qi, G,
                                fuse_seeds ? frank : nullptr);
        } else if (fused) {
            launch_ls_walk_reheat(*g, ctl, b.we, b.wev, G, gray_evals, b.wx, b.wh, x, h, rx, rh, dout, G, dout2,
                                  dbl_evals, c->stream, rt);
        } else {
            launch_ls_after_walk(*g, ctl, b.we, b.wev, G, gray_evals, b.wx, b.wh, x, h, rx, rh, c->stream);
            // post-reheat phase (spin_glass_solver.py:670-717)
            launch_single_best(*g, rx, rh, dout, c->stream);
            launch_ls_reheat_single(*g, ctl, dout, x, h, rx, rh, c->stream);
            launch_double_best(*g, rx, rh, G, p->double_mode, dout2, dsc, c->stream);
            launch_ls_reheat_double(*g, ctl, dout2, x, h, rx, rh, dbl_evals, c->stream);
        }
        return PQI_OK;
    };
    // The iteration is captured once per (graph, parameters, buffers) as one
    // graph, with the read-back of the control record into pinned memory, and
    // replayed: one launch and one synchronisation per iteration.  Events
    // recorded inside a graph cannot be timed here, so with kernel timing the
    // descent is launched directly with its own events and the graph holds the
    // rest of the iteration.  Where graph capture is not available, the two
    // segments between the directly launched descent and walk are captured
    // instead (also selected by PQI_LS_ITER_GRAPH=0).
    if (iter_graph_enabled && !fast && !c->ls.host_ctl) {
        if (hipHostMalloc(&c->ls.host_ctl, sizeof(LsCtl), hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            c->ls.host_ctl = nullptr;
        }
    }
    const bool want_iter = iter_graph_enabled && (fast || c->ls.host_ctl);
    // the start of an iteration: the exact fp64 field of a new round and the
    // iteration's control reset.  In the graph it closes the previous iteration,
    // after the control record has been copied for the host, so that the next
    // iteration starts with the descent
    auto prologue = [&]() {
        launch_field_repair(*g, x, h, ctl, fix, c->stream);
        launch_ls_begin(ctl, c->stream);
    };
    auto tail = [&](int slot) -> int {
        CHK(segment(0, nullptr));
        launch_gray_walk(*g, x, h, b.sel, b.seld, bs, G, T, b.wx, b.wh, b.we, b.wev, c->stream, &ctl->restart);
        if (fast) {
            // the report (in the last fused kernel, or a kernel of its own) also
            // applies the next iteration's control reset (int32 fields are
            // exact: no field repair)
            const LsReportTo rt{rep_dev[slot], xout_dev[slot], c->ls.seq_dev, fix, g->w_up_i};
            CHK(segment(1, fused ? &rt : nullptr));
            if (!fused)
                launch_ls_report_i8(*g, x, h, ctl, fix, rep_dev[slot], xout_dev[slot], c->ls.seq_dev, c->stream);
            return PQI_OK;
        }
        CHK(segment(1, nullptr));
        HIPCHK(hipMemcpyAsync(c->ls.host_ctl, ctl, sizeof(LsCtl), hipMemcpyDeviceToHost, c->stream));
        prologue();
        return PQI_OK;
    };
    // captured once per (graph, parameters, buffers) and replayed per iteration
    // The fast path launches its kernels directly: measured on MI355X, two
    // back-to-back iteration graphs leave about 8 us between them, direct
    // launches queued one iteration ahead none.  PQI_LS_FAST_GRAPH=1 captures
    // and replays the iterations as graphs instead.
    const bool fast_graphs = [] {
        const char* e = getenv("PQI_LS_FAST_GRAPH");
        return e && e[0] == '1';
    }();
    const std::vector<int64_t> key = {(int64_t)g, (int64_t)g->W, (int64_t)g->adj_ptr, n, G, T, bs, (int64_t)timing,
                                      (int64_t)fast_graphs,
                                      (int64_t)want_iter, (int64_t)fix, (int64_t)c->ls.host_ctl, (int64_t)fast,
                                      (int64_t)rep_dev[0], (int64_t)rep_dev[1], (int64_t)c->ls.seq_dev,
                                      p->double_mode, p->seed_engine, p->is_spin_glass, (int64_t)x, (int64_t)h,
                                      (int64_t)rx, (int64_t)rh, (int64_t)dout, (int64_t)dout2, (int64_t)dsc,
                                      (int64_t)ctl, (int64_t)b.wx, (int64_t)b.wh, (int64_t)b.we, (int64_t)b.wev,
                                      (int64_t)b.sel, (int64_t)b.seld, (int64_t)b.delta,
                                      (int64_t)c->bufs[S_BITS].p, (int64_t)c->bufs[S_BATCH].p,
                                      (int64_t)c->bufs[S_RANK].p};
    if (c->ls.key != key) {
        ls_graphs_reset(c);
        // the fast path captures the iteration once per report slot
        for (int slot = 0; want_iter && (fast_graphs || !fast) && slot < (fast ? 2 : 1); ++slot) {
            hipGraphExec_t& exec = slot ? c->ls.iter2 : c->ls.iter;
            hipGraph_t graph = nullptr;
            if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) == hipSuccess) {
                // a failure here only means the capture is not available: the
                // direct launches below repeat the sequence and report real errors
                if (!timing) launch_descent(*g, x, h, cap, dout, c->stream, nullptr, nullptr, &ctl->restart);
                const int r = tail(slot);
                const hipError_t e = hipStreamEndCapture(c->stream, &graph);
                if (r != PQI_OK || e != hipSuccess || !graph ||
                    hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
                    exec = nullptr;
                if (graph) (void)hipGraphDestroy(graph);
            }
            (void)hipGetLastError();
        }
        if (fast && !(c->ls.iter && c->ls.iter2)) ls_graphs_reset(c);  // both slots or neither
        // (the fast path launches directly what it has not captured)
        for (int w = 0; w < 2 && !fast && !c->ls.failed && !c->ls.iter; ++w) {
            hipGraph_t graph = nullptr;
            if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
                c->ls.failed = true;
                break;
            }
            const int r = segment(w, nullptr);
            const hipError_t e = hipStreamEndCapture(c->stream, &graph);
            if (r != PQI_OK) {
                if (graph) (void)hipGraphDestroy(graph);
                return r;
            }
            if (e != hipSuccess || !graph ||
                hipGraphInstantiate(&c->ls.seg[w], graph, nullptr, nullptr, 0) != hipSuccess)
                c->ls.failed = true;
            if (graph) (void)hipGraphDestroy(graph);
        }
        (void)hipGetLastError();
        if (c->ls.failed) ls_graphs_reset(c), c->ls.failed = true;
        c->ls.key = key;
    }
    auto run_segment = [&](int which) -> int {
        if (c->ls.seg[which]) {
            HIPCHK(hipGraphLaunch(c->ls.seg[which], c->stream));
            return PQI_OK;
        }
        return segment(which, nullptr);
    };
    if (fast) {
        // Iteration k + 1 is queued before the host reads iteration k's report
        // (slot k mod 2), so the GPU never waits for the host between
        // iterations.  Once the search has stopped, the iteration reset leaves
        // `restart` set and every phase of the queued iteration does nothing; it
        // finishes on the stream after this call returns.  The host waits for
        // a report by polling its number in pinned memory: no event or marker
        // sits between two iterations on the stream.
        EventTimer dt1(timing);
        EventTimer* dtimer[2] = {&descent_timer, &dt1};
        unsigned long long want[2] = {0, 0};
        c->ls.seq_dirty = true;
        auto enqueue = [&](int slot) -> int {
            hipGraphExec_t exec = slot ? c->ls.iter2 : c->ls.iter;
            if (timing || !exec)
                launch_descent(*g, x, h, cap, dout, c->stream, dtimer[slot]->start_event(),
                               dtimer[slot]->stop_event(), &ctl->restart);
            if (exec) HIPCHK(hipGraphLaunch(exec, c->stream));
            else CHK(tail(slot));
            want[slot] = ++c->ls.seq_host;
            return PQI_OK;
        };
        launch_ls_start_i8(*g, (const uint8_t*)c->pin[0].dev, x, h, rx, rh, ctl, fix, hc.scale, p->max_rounds,
                           p->is_spin_glass, c->stream);
        CHK(enqueue(0));
        int slot = 0;
        for (;;) {
            CHK(enqueue(slot ^ 1));
            CHK(wait_report(c, rep[slot], want[slot]));
            std::memcpy(&hc, &rep[slot]->ctl, sizeof(hc));
            if (timing) {
                st->descent_kernel_seconds += dtimer[slot]->seconds();
                st->descent_launches++;
            }
            if (ls_stopped(&hc)) {
                // The stop, the state and the score are taken only from a
                // report whose check word matches: a number seen before the
                // rest of its report would otherwise leave the slot's previous
                // record (another call's final, stopped one) to be read.  If
                // the whole report goes on, so does the loop.
                CHK(settle_report(c, rep[slot], xout[slot], n, want[slot]));
                std::memcpy(&hc, &rep[slot]->ctl, sizeof(hc));
                if (ls_stopped(&hc)) break;
            }
            slot ^= 1;
        }
        c->ls.seq_dirty = false;
        const LsReport& r = *rep[slot];
        fill_search_stats(c, hc, st);
        field_repair_stats(r.fix, st);
        st->cut_value = r.cut;
        st->min_energy = p->is_spin_glass ? -r.energy : r.energy;
        std::memcpy(theta, xout[slot], (size_t)n);
        return PQI_OK;
    }
    if (c->ls.iter) prologue();
    for (;;) {
        if (!hc.again && p->max_rounds > 0 && hc.rounds >= p->max_rounds) break;
        if (c->ls.iter) {
            if (timing)
                launch_descent(*g, x, h, cap, dout, c->stream, descent_timer.start_event(), descent_timer.stop_event(),
                               &ctl->restart);
            HIPCHK(hipGraphLaunch(c->ls.iter, c->stream));
            CHK(sync(c));
            std::memcpy(&hc, c->ls.host_ctl, sizeof(hc));
            if (timing) {
                st->descent_kernel_seconds += descent_timer.seconds();
                st->descent_launches++;
            }
            if (hc.done) break;
            continue;
        }
        launch_field_repair(*g, x, h, ctl, fix, c->stream);  // a new round: exact fp64 field
        launch_ls_begin(ctl, c->stream);
        launch_descent(*g, x, h, cap, dout, c->stream, descent_timer.start_event(), descent_timer.stop_event());
        CHK(run_segment(0));
        walk_timer.start(c->stream);
        launch_gray_walk(*g, x, h, b.sel, b.seld, bs, G, T, b.wx, b.wh, b.we, b.wev, c->stream, &ctl->restart);
        walk_timer.stop(c->stream);
        CHK(run_segment(1));
        HIPCHK(hipMemcpyAsync(&hc, ctl, sizeof(hc), hipMemcpyDeviceToHost, c->stream));
        CHK(sync(c));
        if (timing) {
            st->descent_kernel_seconds += descent_timer.seconds();
            st->descent_launches++;
            if (hc.walk_ran) {
                st->walk_kernel_seconds += walk_timer.seconds();
                st->walk_launches++;
            }
        }
        if (hc.done) break;
    }
    fill_search_stats(c, hc, st);
    // field repairs, the exact score and the state come back in one
    // synchronisation, through pinned memory
    launch_score_with(*g, x, 1, score, score + 1, hb, c->stream);
    CHK(pin_ensure(c, 0, 64 + (size_t)n));
    char* pb = (char*)c->pin[0].p;
    HIPCHK(hipMemcpyAsync(pb, fix, sizeof(FieldRepair), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(pb + 32, score, 16, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(pb + 64, x, n, hipMemcpyDeviceToHost, c->stream));
    CHK(sync(c));
    FieldRepair f;
    std::memcpy(&f, pb, sizeof(f));
    field_repair_stats(f, st);
    double hv[2];
    std::memcpy(hv, pb + 32, 16);
    st->cut_value = hv[0];
    st->min_energy = p->is_spin_glass ? -hv[1] : hv[1];
    std::memcpy(theta, pb + 64, (size_t)n);
    return PQI_OK;
}

int pqi_local_search(pqi_graph* g, uint8_t* theta, const pqi_search_params* p, pqi_search_stats* st) {
    if (!g || !theta || !st) return fail(PQI_ERR_INVALID, "null argument");
    CHK(check_params(g, p));
    if (g->n < 2) return fail(PQI_ERR_INVALID, "n must be >= 2");
    pqi_ctx* c = g->ctx;
    CHK(set_device(c));
    std::memset(st, 0, sizeof(*st));
    st->learned_min_gain = INFINITY;
    return (p->flags & PQI_FLAG_HOST_LOOP) ? local_search_host(g, theta, p, st) : local_search_device(g, theta, p, st);
}

static int local_search_host(pqi_graph* g, uint8_t* theta, const pqi_search_params* p, pqi_search_stats* st) {
    pqi_ctx* c = g->ctx;
    const int64_t n = g->n;
    const size_t fb = field_bytes(*g);
    const double scale = p->is_spin_glass ? 2.0 : 1.0;
    const int64_t C = n * (n - 1) / 2;
    const int64_t G = p->walkers;
    const int64_t T = (p->gray_iterations + G - 1) / G;

    uint8_t* x;
    void* h;
    CHK(load_state(g, theta, &x, &h));
    FieldRepair* fix;
    CHK(ensure(c, S_FIX, sizeof(FieldRepair), (void**)&fix));
    HIPCHK(hipMemsetAsync(fix, 0, sizeof(FieldRepair), c->stream));
    if (p->flags & PQI_FLAG_TEST_DRIFT) launch_perturb_field(*g, h, c->stream);
    uint8_t* rx;
    void* rh;
    PhaseOut* dout;
    void* dsc;
    CHK(ensure(c, S_RX, n, (void**)&rx));
    CHK(ensure(c, S_RH, n * fb, &rh));
    CHK(ensure(c, S_OUT, sizeof(PhaseOut), (void**)&dout));
    CHK(ensure(c, S_DBL, double_scratch_bytes(*g), &dsc));

    double cut0, e0;
    CHK(score_device(g, x, &cut0, &e0));
    st->max_energy = p->is_spin_glass ? e0 : cut0;

    PhaseOut o;
    std::vector<double> we(G);
    GrayBufs b;
    const bool timing = (p->flags & PQI_FLAG_TIME_KERNELS) != 0;
    EventTimer walk_timer(timing), descent_timer(timing);
    for (;;) {
        if (p->max_rounds > 0 && st->rounds >= p->max_rounds) break;
        st->rounds++;
        launch_field_repair(*g, x, h, nullptr, fix, c->stream);  // a new round: exact fp64 field

        double t0 = now();
        // bounded launches keep every kernel short; repeat while the bound was hit
        const int64_t cap = 16 * n + 4096;
        do {
            launch_descent(*g, x, h, cap, dout, c->stream, descent_timer.start_event(), descent_timer.stop_event());
            CHK(read_out(c, dout, &o));
            if (timing) {
                st->descent_kernel_seconds += descent_timer.seconds();
                st->descent_launches++;
            }
            st->descent_evals += o.evals;
            st->single_evals += o.evals;
            st->single_accepts += o.accepts;
            st->max_energy += scale * o.gain_sum;
        } while (o.accepts == cap);
        st->seconds_single += now() - t0;
        // the descent ends on a failed single-flip check: go on to double flips

        t0 = now();
        launch_double_best(*g, x, h, G, p->double_mode, dout, dsc, c->stream);
        CHK(read_out(c, dout, &o));
        st->double_evals += p->double_mode == PQI_DOUBLE_EXHAUSTIVE ? C : std::min(G, C);
        if (o.gain > 0.0) {
            launch_flip(*g, x, h, o.k, o.l, c->stream);
            st->max_energy += scale * o.gain;
            st->double_accepts++;
            st->seconds_double += now() - t0;
            continue;
        }
        st->seconds_double += now() - t0;

        t0 = now();
        int64_t count;
        CHK(gray_seeds(g, x, h, p, b, &count));
        st->seed_evals += count;
        double top;
        HIPCHK(hipMemcpyAsync(&top, b.seld, 8, hipMemcpyDeviceToHost, c->stream));
        CHK(sync(c));
        if (top > 0.0) {
            CHK(gray_walk(g, x, h, p, b, 1, 0));
            HIPCHK(hipMemcpyAsync(x, b.wx, n, hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(hipMemcpyAsync(h, b.wh, n * fb, hipMemcpyDeviceToDevice, c->stream));
            st->max_energy += scale * top;
            st->seed_accepts++;
            st->seconds_seeds += now() - t0;
            continue;
        }
        st->seconds_seeds += now() - t0;

        t0 = now();
        walk_timer.start(c->stream);
        CHK(gray_walk(g, x, h, p, b, G, T));
        walk_timer.stop(c->stream);
        HIPCHK(hipMemcpyAsync(we.data(), b.we, G * 8, hipMemcpyDeviceToHost, c->stream));
        std::vector<int64_t> wev(G);
        HIPCHK(hipMemcpyAsync(wev.data(), b.wev, G * 8, hipMemcpyDeviceToHost, c->stream));
        CHK(sync(c));
        st->walk_kernel_seconds += walk_timer.seconds();
        st->walk_launches += timing ? 1 : 0;
        for (int64_t w = 0; w < G; ++w) st->gray_evals_computed += wev[w];
        st->gray_evals += G * reference_gray_steps(T, n);
        int64_t bw = 0;
        for (int64_t w = 1; w < G; ++w)
            if (we[w] > we[bw]) bw = w;
        const double walk = we[bw];
        st->seconds_walk += now() - t0;
        if (walk > 0.0) {
            HIPCHK(hipMemcpyAsync(x, b.wx + bw * n, n, hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(hipMemcpyAsync(h, (char*)b.wh + bw * n * fb, n * fb, hipMemcpyDeviceToDevice, c->stream));
            st->max_energy += scale * walk;
            st->walk_accepts++;
            continue;
        }

        // post-reheat phase (spin_glass_solver.py:670-717)
        t0 = now();
        HIPCHK(hipMemcpyAsync(rx, b.wx + bw * n, n, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(rh, (char*)b.wh + bw * n * fb, n * fb, hipMemcpyDeviceToDevice, c->stream));
        double reheat = walk;
        launch_single_best(*g, rx, rh, dout, c->stream);
        CHK(read_out(c, dout, &o));
        st->single_evals += o.evals;
        if (o.gain + reheat > 0.0) {
            HIPCHK(hipMemcpyAsync(x, rx, n, hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(hipMemcpyAsync(h, rh, n * fb, hipMemcpyDeviceToDevice, c->stream));
            launch_flip(*g, x, h, o.k, -1, c->stream);
            st->max_energy += scale * (reheat + o.gain);
            st->reheat_accepts++;
            st->seconds_reheat += now() - t0;
            continue;
        }
        if (o.gain > 0.0) {
            launch_flip(*g, rx, rh, o.k, -1, c->stream);
            reheat += o.gain;
        }
        launch_double_best(*g, rx, rh, G, p->double_mode, dout, dsc, c->stream);
        CHK(read_out(c, dout, &o));
        st->double_evals += p->double_mode == PQI_DOUBLE_EXHAUSTIVE ? C : std::min(G, C);
        if (o.gain + reheat > 0.0) {
            HIPCHK(hipMemcpyAsync(x, rx, n, hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(hipMemcpyAsync(h, rh, n * fb, hipMemcpyDeviceToDevice, c->stream));
            launch_flip(*g, x, h, o.k, o.l, c->stream);
            st->max_energy += scale * (reheat + o.gain);
            st->reheat_accepts++;
            st->seconds_reheat += now() - t0;
            continue;
        }
        st->seconds_reheat += now() - t0;
        break;
    }
    double cut, en;
    CHK(read_field_repair(c, fix, st));
    CHK(score_device(g, x, &cut, &en));
    st->cut_value = cut;
    st->min_energy = p->is_spin_glass ? -en : en;
    HIPCHK(hipMemcpyAsync(theta, x, n, hipMemcpyDeviceToHost, c->stream));
    return sync(c);
}

}  // extern "C"
