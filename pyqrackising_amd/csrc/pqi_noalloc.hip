// pqi_eval_batch_noalloc: candidate scoring that allocates no device memory.
//
// The same work as pqi_eval_batch_ex with the exact engine (staging of the
// caller's packed rows, the contraction, the values and the first argmax),
// with every device buffer it needs carved out of one file-scope array of
// kScratchBytes.  Each GPU the library runs on holds its own copy of that
// array from the moment the library's code object is loaded on it until the
// process ends.  On the host, `scratch` names a host-side stand-in: its device
// address on each GPU is taken once, by pqi_create with that GPU current, and
// only that address is handed to kernels and copies.
//
// Every call on one GPU shares that GPU's array and leaves its contents
// behind; whatever must start at zero is cleared by the kernels of each call.
// Calls that overlap on one GPU (two contexts on different streams, or two
// host threads) race on the array.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "pqi_internal.h"

namespace pqi {

constexpr size_t kScratchBytes = (size_t)32 << 20;  // 32 MiB per GPU
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_mu;
static void* g_ptr[kMaxDevices];

int noalloc_init(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(PQI_ERR_UNSUPPORTED, "device index above 63");
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device]) return PQI_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
        return fail(PQI_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_ptr[device] = p;
    return PQI_OK;
}

static void* scratch_of(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    return g_ptr[device];
}

struct BestOut {
    int64_t index;
    double value;
};

// values[b] from q[b] = s_b^T W s_b exactly as pqi_eval_batch_ex forms them on
// the host (integer arithmetic for int8 weights), and the first maximum.
__global__ __launch_bounds__(1024) void k_values_best(const double* __restrict__ q, int64_t B, int is_i8,
                                                      int is_spin_glass, long long w_up_i, double w_up,
                                                      double* __restrict__ vals, BestOut* best) {
    __shared__ double sv[16];
    __shared__ int64_t sk[16];
    double bv = -INFINITY;
    int64_t bk = INT64_MAX;
    for (int64_t b = threadIdx.x; b < B; b += blockDim.x) {
        const double qv = q[b];
        double v;
        if (is_i8) {
            const long long qi = (long long)qv;
            v = is_spin_glass ? (double)(-qi / 2) : (double)((2 * w_up_i - qi) / 4);
        } else {
            v = is_spin_glass ? -qv / 2.0 : (2.0 * w_up - qv) / 4.0;
        }
        vals[b] = v;
        if (v > bv || (v == bv && b < bk)) {
            bv = v;
            bk = b;
        }
    }
    // wave then block first-argmax (value, then lowest index)
    for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_down(bv, off, 64);
        const int64_t ok = __shfl_down(bk, off, 64);
        if (ov > bv || (ov == bv && ok < bk)) {
            bv = ov;
            bk = ok;
        }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        sv[wid] = bv;
        sk[wid] = bk;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int w = 1; w < nw; ++w)
            if (sv[w] > bv || (sv[w] == bv && sk[w] < bk)) {
                bv = sv[w];
                bk = sk[w];
            }
        best->index = bk;
        best->value = bv;
    }
}

static size_t up256(size_t v) { return (v + 255) & ~(size_t)255; }

}  // namespace pqi

using namespace pqi;

extern "C" int pqi_eval_batch_noalloc(pqi_graph* g, const uint32_t* bits, int64_t B, int is_spin_glass,
                                      double* values, int64_t* best_index, double* best_value) {
    if (!g || !bits || !best_index || !best_value) return fail(PQI_ERR_INVALID, "null argument");
    if (B < 1) return fail(PQI_ERR_INVALID, "B must be >= 1");
    pqi_ctx* c = g->ctx;
    const int64_t nw = (g->n + 31) >> 5;
    const int64_t wpr = g->n_pad >> 5;
    // layout in the scratch array: rows at the device pitch, host rows staged
    // when the pitches differ, row totals, contraction partials, values, best
    const size_t s_bits = up256((size_t)B * wpr * 4);
    const size_t s_stage = nw != wpr ? up256((size_t)B * nw * 4) : 0;
    const size_t s_q = up256((size_t)B * 8);
    const size_t s_part = up256(batch_scratch_bytes(*g, B));
    const size_t s_vals = up256((size_t)B * 8);
    const size_t need = s_bits + s_stage + s_q + s_part + s_vals + up256(sizeof(BestOut));
    if (need > kScratchBytes)
        return fail(PQI_ERR_INVALID, "pqi_eval_batch_noalloc: B = " + std::to_string(B) + " at n = " +
                                         std::to_string(g->n) + " needs " + std::to_string(need) +
                                         " bytes of scratch, above its maximum of " + std::to_string(kScratchBytes) +
                                         "; use pqi_eval_batch_ex");
    if (hipSetDevice(c->device) != hipSuccess) return fail(PQI_ERR_HIP, "hipSetDevice failed");
    unsigned char* base = (unsigned char*)scratch_of(c->device);
    if (!base) return fail(PQI_ERR_HIP, "the scratch address of this device was not taken (pqi_create failed?)");
    uint32_t* dbits = (uint32_t*)base;
    uint32_t* stage = (uint32_t*)(base + s_bits);
    double* q = (double*)(base + s_bits + s_stage);
    void* part = base + s_bits + s_stage + s_q;
    double* vals = (double*)(base + s_bits + s_stage + s_q + s_part);
    BestOut* best = (BestOut*)(base + s_bits + s_stage + s_q + s_part + s_vals);
    hipStream_t s = c->stream;
    hipError_t e;
    if (nw == wpr) {
        e = hipMemcpyAsync(dbits, bits, (size_t)B * nw * 4, hipMemcpyHostToDevice, s);
    } else {
        // padding words become zero; the padded columns of W are zero too
        e = hipMemcpyAsync(stage, bits, (size_t)B * nw * 4, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) launch_restride_words(dbits, wpr, stage, nw, B, s);
    }
    if (e != hipSuccess) return fail(PQI_ERR_HIP, std::string("upload of the rows: ") + hipGetErrorString(e));
    launch_batch_quadratic(*g, dbits, wpr, B, q, part, s);
    hipLaunchKernelGGL(k_values_best, dim3(1), dim3(1024), 0, s, q, B, g->kind == PQI_WEIGHTS_I8 ? 1 : 0,
                       is_spin_glass ? 1 : 0, (long long)g->w_up_i, g->w_up, vals, best);
    e = hipGetLastError();
    if (e != hipSuccess) return fail(PQI_ERR_HIP, std::string("launch: ") + hipGetErrorString(e));
    BestOut hb;
    if (values) e = hipMemcpyAsync(values, vals, (size_t)B * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(&hb, best, sizeof(hb), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail(PQI_ERR_HIP, std::string("read-back: ") + hipGetErrorString(e));
    *best_index = hb.index;
    *best_value = hb.value;
    return PQI_OK;
}
