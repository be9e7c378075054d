// Batched candidate energies q_b = s_b^T W s_b for gfx950.
//
// Candidates arrive packed (B rows of uint32 words, LSB-first, the layout of
// pyqrackising/maxcut_tfim_util.py:304-316) and are unpacked to +-1 in LDS.
// The contraction P = S W runs on MFMA and the epilogue forms
// q_b = sum_j S_bj P_bj without writing P: this is the dense contraction
// E = rowsum((S W) .* S) that replaces kernels.cl:44-94 (calculate_cut), which
// walks the upper triangle of W once per candidate.
//
//   int8 weights: v_mfma_i32_32x32x32_i8, int32 accumulation, int64 row totals
//   (exact); fp32/fp64 weights: FP64 tiles (exact to FP64 rounding).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pqi_internal.h"
#include "pqi_tile_i8.h"

namespace pqi {

__global__ void k_batch_sum(const double* part, int64_t nblk, int64_t B, double* q);


namespace {
constexpr int BM = TILE_BM;  // candidates per work group
constexpr int BN = TILE_BN;  // output columns per work group
}  // namespace

// q[m] += sum over this block's 128 columns j of s_mj (S W)_mj.  q is zeroed by the caller.
__global__ __launch_bounds__(256) void k_batch_i8(const int8_t* __restrict__ W, int64_t ld, int64_t n_pad,
                                                  const uint32_t* __restrict__ bits, int64_t wpr, int64_t B,
                                                  unsigned long long* __restrict__ q, const int32_t* gate) {
    if (gate && *gate) return;
    __shared__ __align__(16) int8_t As[TILE_BM * (64 + 16)];
    __shared__ __align__(16) int8_t Bs[TILE_BN * (64 + 16)];
    __shared__ uint32_t Sb[TILE_BM * (TILE_BN / 32)];
    tile_i8<64>(W, ld, n_pad, bits, wpr, B, (int64_t)blockIdx.y * TILE_BM, (int64_t)blockIdx.x * TILE_BN, q, As, Bs,
                Sb);
}

// Wide tiles, XCD-aware order: block b runs on XCD b % 8 (round-robin
// dispatch); the candidate tiles of one column block get the same residue and
// consecutive slots, so they share that column block of W through one L2.
// The grid is rounded up to whole groups of 8 column blocks; the extra blocks
// exit at once.
// grid: 8-column-block groups x MB candidate blocks x KS splits of K
template <int BK, int NW>
__global__ __launch_bounds__(64 * NW) void k_batch_i8_wide(const int8_t* __restrict__ W, int64_t ld, int64_t n_pad,
                                                       const uint32_t* __restrict__ bits, int64_t wpr, int64_t B,
                                                       unsigned long long* __restrict__ q, int64_t MB, int64_t JB,
                                                       int64_t KS, const int32_t* gate) {
    extern __shared__ __align__(16) unsigned char lds[];
    if (gate && *gate) return;
    const int64_t b = blockIdx.x, x = b & 7, slot = b >> 3;
    const int64_t ks = slot % KS, t = slot / KS;
    const int64_t jb = (t / MB) * 8 + x, mb = t % MB;
    if (jb >= JB) return;
    const int64_t kl = n_pad / KS;
    tile_i8_wide<BK, NW>(W, ld, ks * kl, (ks + 1) * kl, bits, wpr, B, mb * WIDE_BM, jb * WIDE_BN, q, lds);
}

// ---------------------------------------------------------------------------
// Split-BF16 variant for fp32 weights: W = W_hi + W_lo + r with
// W_hi = bf16(W), W_lo = bf16(W - W_hi), |r| <= 2^-18 |W|.  The weights are
// stored as one n_pad x 2 n_pad bf16 matrix [W_hi | W_lo]; the contraction runs
// over K = 2 n_pad with the candidate repeated, on v_mfma_f32_32x32x16_bf16
// with FP32 accumulation.  +-1 is exact in bf16.  Row partial sums per wave
// column block go to part[] and are added in FP64 in fixed order.
// ---------------------------------------------------------------------------

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

// round-to-nearest-even fp32 -> bf16 bits (finite inputs)
__device__ __forceinline__ uint16_t bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__global__ void k_split_bf16(const float* __restrict__ W, int64_t n_pad, uint16_t* __restrict__ Wcat) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_pad * n_pad) return;
    const int64_t j = t / n_pad, k = t % n_pad;
    const float w = W[t];
    const uint16_t hi = bf16_rne(w);
    const uint16_t lo = bf16_rne(w - bf16_to_f(hi));
    Wcat[j * 2 * n_pad + k] = hi;
    Wcat[j * 2 * n_pad + n_pad + k] = lo;
}

void launch_split_bf16(const float* W, int64_t n_pad, uint16_t* Wcat, hipStream_t s) {
    const int64_t tot = n_pad * n_pad;
    hipLaunchKernelGGL(k_split_bf16, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, W, n_pad, Wcat);
}

__global__ __launch_bounds__(256) void k_batch_bf16(const uint16_t* __restrict__ Wcat, int64_t n_pad,
                                                    const uint32_t* __restrict__ bits, int64_t wpr, int64_t B,
                                                    double* __restrict__ part, int64_t nparts) {
    constexpr int BKB = 32;              // bf16 elements per stage
    constexpr int ROWB = BKB * 2 + 16;   // padded LDS row in bytes
    __shared__ __align__(16) unsigned char As[BM * ROWB];
    __shared__ __align__(16) unsigned char Bs[BN * ROWB];
    __shared__ uint32_t Sb[BM * (BN / 32)];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm = wid >> 1, wn = wid & 1;
    const int64_t m0 = (int64_t)blockIdx.y * BM, j0 = (int64_t)blockIdx.x * BN;
    const int lr = tid >> 1, lh = tid & 1;
    const int64_t K = 2 * n_pad;

    v16f acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.f;

    const int64_t am = m0 + lr;
    for (int64_t k0 = 0; k0 < K; k0 += BKB) {
        const int64_t kk = k0 < n_pad ? k0 : k0 - n_pad;
        {
            // each of the two loader threads of a row unpacks 16 of the 32 bits
            const uint32_t word = (am < B) ? bits[am * wpr + (kk >> 5)] : 0u;
            const uint32_t half = (word >> (16 * lh)) & 0xFFFFu;
            uint32_t o[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const uint32_t b0 = (half >> (2 * t)) & 1u, b1 = (half >> (2 * t + 1)) & 1u;
                o[t] = (0x3F80u | (b0 << 15)) | ((0x3F80u | (b1 << 15)) << 16);
            }
            uint4* dst = (uint4*)&As[lr * ROWB + lh * 32];
            dst[0] = make_uint4(o[0], o[1], o[2], o[3]);
            dst[1] = make_uint4(o[4], o[5], o[6], o[7]);
        }
        {
            const uint4* src = (const uint4*)(Wcat + (j0 + lr) * K + k0 + lh * 16);
            uint4 a = src[0], b = src[1];
            uint4* dst = (uint4*)&Bs[lr * ROWB + lh * 32];
            dst[0] = a;
            dst[1] = b;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < BKB; ks += 16) {
            v8bf av[2], bv[2];
#pragma unroll
            for (int mi = 0; mi < 2; ++mi)
                av[mi] = *(const v8bf*)&As[(wm * 64 + mi * 32 + (lane & 31)) * ROWB + (ks + 8 * (lane >> 5)) * 2];
#pragma unroll
            for (int nj = 0; nj < 2; ++nj)
                bv[nj] = *(const v8bf*)&Bs[(wn * 64 + nj * 32 + (lane & 31)) * ROWB + (ks + 8 * (lane >> 5)) * 2];
#pragma unroll
            for (int mi = 0; mi < 2; ++mi)
#pragma unroll
                for (int nj = 0; nj < 2; ++nj)
                    acc[mi][nj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[mi], bv[nj], acc[mi][nj], 0, 0, 0);
        }
        __syncthreads();
    }
    for (int t = tid; t < BM * (BN / 32); t += 256) {
        const int r = t >> 2, c = t & 3;
        const int64_t m = m0 + r;
        Sb[t] = (m < B) ? bits[m * wpr + (j0 >> 5) + c] : 0u;
    }
    __syncthreads();
    float vals[32];
#pragma unroll
    for (int i = 0; i < 32; ++i) vals[i] = 0.f;
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int nj = 0; nj < 2; ++nj) {
            const int col = wn * 64 + nj * 32 + (lane & 31);
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = wm * 64 + mi * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                const uint32_t bit = (Sb[row * 4 + (col >> 5)] >> (col & 31)) & 1u;
                const float p = acc[mi][nj][r];
                vals[mi * 16 + r] += bit ? -p : p;
            }
        }
#pragma unroll
    for (int t = 0, half = 16; t < 5; ++t, half >>= 1) {
        const int up = (lane >> t) & 1;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            if (i < half) {
                const float send = up ? vals[i] : vals[i + half];
                const float keep = up ? vals[i + half] : vals[i];
                vals[i] = keep + __shfl_xor(send, 1 << t, 64);
            }
        }
    }
    const int qi = (lane & 1) * 16 + ((lane >> 1) & 1) * 8 + ((lane >> 2) & 1) * 4 + ((lane >> 3) & 1) * 2 +
                   ((lane >> 4) & 1);
    const int mi = qi >> 4, r = qi & 15;
    const int row = wm * 64 + mi * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    const int64_t m = m0 + row;
    if (m < B) part[m * nparts + blockIdx.x * 2 + wn] = (double)vals[0];
}

// ---------------------------------------------------------------------------
// Wide split-BF16 tile: 256 candidates x 256 output columns per work group of
// 8 waves (wave w: candidate rows [64 (w mod 4), +64), column blocks
// [4 (w / 4), +4); 128 fp32 accumulators per lane, two waves per SIMD).  The
// contraction runs over K = 2 n_pad (W_hi then W_lo, the candidate's spins
// twice) in chunks of 64 bf16: a W_cat chunk is 128 bytes per row, the same
// LDS layout and loaders as the 128-deep int8 wide tile.  The +-1 A
// fragments are made in registers from the candidate bits (sign bit over
// 1.0).  Each wave writes its rows' partial sums (its 128 columns) to
// part[m * 2 JB + 2 jb + column half]; they are summed in a fixed order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ v8bf unpack8_bf16(uint32_t x) {
    v4i w;
#pragma unroll
    for (int u = 0; u < 4; ++u)
        w[u] = (int)(0x3F803F80u | ((x << (15 - 2 * u)) & 0x8000u) | ((x << (30 - 2 * u)) & 0x80000000u));
    return __builtin_bit_cast(v8bf, w);
}

constexpr int BF_KC = 64;                  // bf16 per chunk
constexpr int BF_ROW = 2 * BF_KC + 16;     // padded LDS row in bytes
constexpr size_t bf16_wide_lds_bytes() { return 2 * (size_t)WIDE_BN * BF_ROW + (size_t)WIDE_BM * 8 * 4; }

template <int NCB>
__device__ __forceinline__ void wide_compute_bf16(v16f (&acc)[2][NCB], const uint32_t (&aw)[2][BF_KC / 32],
                                                  const int8_t* Ws, int lc, int hi, int cb0) {
#pragma unroll
    for (int t = 0; t < BF_KC / 16; ++t) {
        v8bf av[2];
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) av[rb] = unpack8_bf16((aw[rb][t >> 1] >> (16 * (t & 1) + 8 * hi)) & 0xFFu);
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) {
            const v8bf bv = *(const v8bf*)&Ws[((cb0 + cb) * 32 + lc) * BF_ROW + 32 * t + 16 * hi];
#pragma unroll
            for (int rb = 0; rb < 2; ++rb)
                acc[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[rb], bv, acc[rb][cb], 0, 0, 0);
        }
    }
}

// grid: XCD-aware as k_batch_i8_wide (8-column-block groups x MB candidate blocks)
__global__ __launch_bounds__(512) void k_batch_bf16_wide(const uint16_t* __restrict__ Wcat, int64_t n_pad,
                                                         const uint32_t* __restrict__ bits, int64_t wpr, int64_t B,
                                                         double* __restrict__ part, int64_t MB, int64_t JB) {
    constexpr int NW = 8, NT = 64 * NW, NCB = 32 / NW;
    constexpr int CB = 2 * BF_KC;                 // bytes per W_cat row per chunk
    constexpr int STAGE = WIDE_BN * BF_ROW;
    constexpr int LPR = CB / 16, RPP = NT / LPR, NP = WIDE_BN / RPP;
    extern __shared__ __align__(16) unsigned char lds[];
    const int64_t blk = blockIdx.x, xcd = blk & 7, slot = blk >> 3;
    const int64_t jb = (slot / MB) * 8 + xcd, mb = slot % MB;
    if (jb >= JB) return;  // whole work group, before any barrier
    const int64_t m0 = mb * WIDE_BM, j0 = jb * WIDE_BN;
    int8_t* stage0 = (int8_t*)lds;
    int8_t* stage1 = stage0 + STAGE;
    uint32_t* Sb = (uint32_t*)(lds + 2 * STAGE);
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int hi = lane >> 5, lc = lane & 31;
    const int rg = wid & 3, ch = wid >> 2, cb0 = ch * NCB;
    const uint32_t* arow[2];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
        const int64_t m = m0 + rg * 64 + rb * 32 + lc;
        arow[rb] = bits + (m < B ? m : 0) * wpr;
    }
    const int64_t pitch = 4 * n_pad;  // bytes per W_cat row (2 n_pad bf16)
    const int8_t* Wb = (const int8_t*)Wcat;
    const int8_t* wbase[NP];
    int woff[NP];
#pragma unroll
    for (int t = 0; t < NP; ++t) {
        const int r = t * RPP + tid / LPR, piece = tid % LPR;
        const int64_t jrow = j0 + r < n_pad ? j0 + r : n_pad - 1;  // rows past n_pad: a valid row, dropped later
        wbase[t] = Wb + jrow * pitch + 16 * piece;
        woff[t] = r * BF_ROW + 16 * piece;
    }
    v16f acc[2][NCB];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < NCB; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.f;

    const int64_t nc = 2 * n_pad / BF_KC;
    auto cl = [nc](int64_t c) { return c < nc ? c : nc - 1; };           // past the end: the last chunk
    auto boff = [&](int64_t c) { return cl(c) * CB; };                   // byte offset in a W_cat row
    auto kbit = [&](int64_t c) { return (cl(c) * BF_KC) % n_pad; };     // the chunk's spin index
    uint32_t a0[2][BF_KC / 32], a1[2][BF_KC / 32];
    v4i w0[NP], w1[NP];
    wide_load_a<BF_KC>(a0, arow, kbit(0));
    wide_load_w<NP>(w0, wbase, boff(0));
    wide_load_w<NP>(w1, wbase, boff(1));
    wide_store_w<NP>(w0, stage0, woff);
    wide_load_w<NP>(w0, wbase, boff(2));
    __syncthreads();
    int64_t c = 0;
    for (; c + 1 < nc; c += 2) {
        wide_load_a<BF_KC>(a1, arow, kbit(c + 1));
        wide_compute_bf16<NCB>(acc, a0, stage0, lc, hi, cb0);
        wide_store_w<NP>(w1, stage1, woff);
        wide_load_w<NP>(w1, wbase, boff(c + 3));
        __syncthreads();
        wide_load_a<BF_KC>(a0, arow, kbit(c + 2));
        wide_compute_bf16<NCB>(acc, a1, stage1, lc, hi, cb0);
        wide_store_w<NP>(w0, stage0, woff);
        wide_load_w<NP>(w0, wbase, boff(c + 4));
        __syncthreads();
    }
    if (c < nc) wide_compute_bf16<NCB>(acc, a0, stage0, lc, hi, cb0);

    const int64_t cwords = (n_pad - j0) >> 5;
    for (int t = tid; t < WIDE_BM * 8; t += NT) {
        const int r = t >> 3, cc = t & 7;
        const int64_t m = m0 + r;
        Sb[t] = (m < B && cc < cwords) ? bits[m * wpr + (j0 >> 5) + cc] : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb)
        if (j0 + (cb0 + cb) * 32 >= n_pad)
#pragma unroll
            for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[rb][cb][r] = 0.f;
    float vals[32];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = rg * 64 + rb * 32 + 8 * (r >> 2) + 4 * hi + (r & 3);
            const v4i wv = *(const v4i*)&Sb[row * 8 + 4 * ch];  // this wave's 4 column words
            float v = 0.f;
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb) {
                const float p = acc[rb][cb][r];
                v += ((wv[cb] >> lc) & 1) ? -p : p;
            }
            vals[rb * 16 + r] = v;
        }
#pragma unroll
    for (int t = 0, half = 16; t < 5; ++t, half >>= 1) {
        const int up = (lane >> t) & 1;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            if (i < half) {
                const float send = up ? vals[i] : vals[i + half];
                const float keep = up ? vals[i + half] : vals[i];
                vals[i] = keep + __shfl_xor(send, 1 << t, 64);
            }
        }
    }
    const int qi = (lane & 1) * 16 + ((lane >> 1) & 1) * 8 + ((lane >> 2) & 1) * 4 + ((lane >> 3) & 1) * 2 +
                   ((lane >> 4) & 1);
    const int rb = qi >> 4, r = qi & 15;
    const int64_t m = m0 + rg * 64 + rb * 32 + 8 * (r >> 2) + 4 * hi + (r & 3);
    if (m < B) part[m * (2 * JB) + 2 * jb + ch] = (double)vals[0];
}

void launch_batch_bf16(const Graph& g, const uint32_t* bits, int64_t wpr, int64_t B, double* q, void* scratch,
                       hipStream_t s) {
    double* part = (double*)scratch;
    static const bool wide_ok = hipFuncSetAttribute((const void*)k_batch_bf16_wide,
                                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                                    (int)bf16_wide_lds_bytes()) == hipSuccess;
    if (B >= 512 && wide_ok) {
        const int64_t MB = (B + WIDE_BM - 1) / WIDE_BM, JB = (g.n_pad + WIDE_BN - 1) / WIDE_BN;
        const int64_t blocks = ((JB + 7) / 8) * 8 * MB;
        hipLaunchKernelGGL(k_batch_bf16_wide, dim3((unsigned)blocks), dim3(512), bf16_wide_lds_bytes(), s,
                           (const uint16_t*)g.Wbf, g.n_pad, bits, wpr, B, part, MB, JB);
        hipLaunchKernelGGL(k_batch_sum, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, s, part, 2 * JB, B, q);
        return;
    }
    const int64_t nparts = 2 * (g.n_pad / BN);
    dim3 grid((unsigned)(g.n_pad / BN), (unsigned)((B + BM - 1) / BM));
    hipLaunchKernelGGL(k_batch_bf16, grid, dim3(256), 0, s, (const uint16_t*)g.Wbf, g.n_pad, bits, wpr, B, part,
                       nparts);
    hipLaunchKernelGGL(k_batch_sum, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, s, part, nparts, B, q);
}

size_t batch_bf16_scratch_bytes(const Graph& g, int64_t B) {
    const int64_t parts = std::max<int64_t>(2 * (g.n_pad / BN), 2 * ((g.n_pad + WIDE_BN - 1) / WIDE_BN));
    return (size_t)B * parts * 8;
}

__global__ void k_gather_rows(const uint32_t* bits, int64_t wpr, const int64_t* sel, int64_t k, uint32_t* out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= k * wpr) return;
    const int64_t r = t / wpr, w = t % wpr;
    out[t] = bits[sel[r] * wpr + w];
}

void launch_gather_rows(const uint32_t* bits, int64_t wpr, const int64_t* sel, int64_t k, uint32_t* out,
                        hipStream_t s) {
    hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((k * wpr + 255) / 256)), dim3(256), 0, s, bits, wpr, sel, k,
                       out);
}

// FP64 variant for real weights: tile 64 candidates x 64 columns, 4 x 4 per
// thread; partial row sums per column block go to part[B][nblk] and are added
// in fixed order by k_batch_sum (deterministic).
template <typename WT>
__global__ __launch_bounds__(256) void k_batch_real(const WT* __restrict__ W, int64_t ld, int64_t n,
                                                    const uint32_t* __restrict__ bits, int64_t wpr, int64_t B,
                                                    double* __restrict__ part, int64_t nblk) {
    constexpr int TB = 64, TK = 16;
    __shared__ double Ss[TK][TB + 1];
    __shared__ double Ws[TK][TB + 1];
    __shared__ double red[TB][17];
    const int tid = threadIdx.x;
    const int tr = tid >> 4, tc = tid & 15;  // 16 x 16 threads, 4 x 4 outputs each
    const int64_t m0 = (int64_t)blockIdx.y * TB, j0 = (int64_t)blockIdx.x * TB;
    double acc[4][4] = {};
    for (int64_t i0 = 0; i0 < n; i0 += TK) {
        for (int t = tid; t < TB * TK; t += 256) {
            const int c = t / TK, i = t % TK;  // candidate c, contraction index i
            const int64_t m = m0 + c, ii = i0 + i;
            double sv = 0.0;
            if (m < B && ii < n) sv = ((bits[m * wpr + (ii >> 5)] >> (ii & 31)) & 1u) ? -1.0 : 1.0;
            Ss[i][c] = sv;
            const int i2 = t / TB, jj = t % TB;
            const int64_t wi = i0 + i2, wj = j0 + jj;
            Ws[i2][jj] = (wi < n && wj < n) ? (double)W[wi * ld + wj] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < TK; ++i) {
            double a[4], b[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) a[u] = Ss[i][tr * 4 + u];
#pragma unroll
            for (int u = 0; u < 4; ++u) b[u] = Ws[i][tc * 4 + u];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[u][v] = fma(a[u], b[v], acc[u][v]);
        }
        __syncthreads();
    }
    // epilogue: multiply by s_mj and sum this thread's 4 columns
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int64_t m = m0 + tr * 4 + u;
        double sacc = 0.0;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int64_t j = j0 + tc * 4 + v;
            if (m < B && j < n) {
                const bool bit = (bits[m * wpr + (j >> 5)] >> (j & 31)) & 1u;
                sacc += bit ? -acc[u][v] : acc[u][v];
            }
        }
        red[tr * 4 + u][tc] = sacc;
    }
    __syncthreads();
    if (tid < TB) {
        double s = 0.0;
        for (int c = 0; c < 16; ++c) s += red[tid][c];
        const int64_t m = m0 + tid;
        if (m < B) part[m * nblk + blockIdx.x] = s;
    }
}

// FP64 matrix-core re-score of a few candidates (C <= 16), real weights:
// H = W S (rows x candidates) on v_mfma_f64_16x16x4f64, then
// q_c = sum_i s_ci H_ic; W is streamed once.  A wave owns a block of 16 rows
// and a split of the columns; per 32 columns each lane loads 8 consecutive
// weights of row (lane mod 16) (the row's four lanes read whole 128-byte
// lines) and the 8 spins of candidate (lane mod 16) at the same columns, and
// feeds 8 MFMAs (A[row][k] = W, B[k][cand] = +-1, k = lane / 16).  Rows and columns up to
// n_pad are read without a branch: W is zero there.  D layout (measured on
// gfx950): lane l, register r -> row 4 r + l / 16, candidate l mod 16.  Each
// wave writes one partial per candidate; k_sum_parts adds them in a fixed
// order.
constexpr int kFewMax = 16;
typedef double v4d __attribute__((ext_vector_type(4)));

template <typename WT>
__global__ __launch_bounds__(256) void k_rescore_mfma(const WT* __restrict__ W, int64_t ld, int64_t nrb,
                                                      const uint32_t* __restrict__ bits, int64_t wpr, int C,
                                                      int64_t ncs, int64_t cols_per, double* __restrict__ part) {
    const int lane = threadIdx.x & 63;
    const int64_t gw = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (gw >= nrb * ncs) return;  // whole wave; no barrier in this kernel
    const int64_t rb = gw / ncs, cs = gw - rb * ncs;
    const int64_t i0 = rb * 16, jb = cs * cols_per, je = jb + cols_per < ld ? jb + cols_per : ld;
    const int r16 = lane & 15, q = lane >> 4;
    const int c = r16;  // this lane's candidate column of B and D
    const bool cl = c < C;
    // 32 columns per step: the four lanes of a row read 8 consecutive
    // weights each, so every step reads whole 128-byte lines of the row;
    // MFMA u takes columns j0 + 8 q + u (q = lane >> 4)
    const WT* Wr = W + (i0 + r16) * ld + 8 * q;
    const uint32_t* brow = bits + (cl ? c : 0) * wpr;
    v4d acc = {0.0, 0.0, 0.0, 0.0};
    for (int64_t j0 = jb; j0 < je; j0 += 32) {
        const uint32_t byte = cl ? (brow[j0 >> 5] >> (8 * q)) & 0xFFu : 0u;
        double a[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) a[u] = (double)Wr[j0 + u];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const double b = cl ? (((byte >> u) & 1u) ? -1.0 : 1.0) : 0.0;
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b, acc, 0, 0, 0);
        }
    }
    // q_c partial: the spins of candidate c at this lane's rows 4 r + q
    const uint32_t rw = cl ? brow[i0 >> 5] >> ((i0 & 31) + q) : 0u;
    double v = 0.0;
#pragma unroll
    for (int r = 0; r < 4; ++r) v += ((rw >> (4 * r)) & 1u) ? -acc[r] : acc[r];
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    if (lane < 16 && cl) part[c * (nrb * ncs) + gw] = v;
}

// the FP64-MFMA grid: 16-row blocks x column splits (multiples of 32, at
// least 256 columns each), about 8192 waves in all
static void mfma_grid(const Graph& g, int64_t* nrb, int64_t* ncs, int64_t* cols_per) {
    *nrb = (g.n + 15) / 16;
    const int64_t want = 8192;
    *ncs = std::max<int64_t>(1, std::min<int64_t>((want + *nrb - 1) / *nrb, g.n_pad / 256));
    *cols_per = ((g.n_pad + *ncs - 1) / *ncs + 31) / 32 * 32;
}

// q[m] = sum of part[m * nparts + p] over p: one work group per candidate,
// strided partial sums then a fixed-order tree (deterministic)
__global__ __launch_bounds__(256) void k_sum_parts(const double* __restrict__ part, int64_t nparts, double* q) {
    __shared__ double red[256];
    const int64_t m = blockIdx.x;
    double s = 0.0;
    for (int64_t p = threadIdx.x; p < nparts; p += 256) s += part[m * nparts + p];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) q[m] = red[0];
}

__global__ void k_batch_sum(const double* part, int64_t nblk, int64_t B, double* q) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= B) return;
    double s = 0.0;
    for (int64_t b = 0; b < nblk; ++b) s += part[m * nblk + b];
    q[m] = s;
}

__global__ void k_zero_u64(unsigned long long* p, int64_t count) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0ull;
}

__global__ void k_i64_to_f64(const unsigned long long* qi, int64_t B, double* q) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m < B) q[m] = (double)(long long)qi[m];
}

// CSR graphs: q_b = sum_u s_u sum_{e in adj(u)} w_e s_{v_e}.  One work group per
// candidate; the candidate's bit row is staged in LDS, rows are strided over
// threads, and the block total is formed in a fixed order (deterministic).
template <typename WT, typename AT>
__global__ __launch_bounds__(256) void k_batch_csr(const int32_t* __restrict__ ap, const int32_t* __restrict__ ai,
                                                   const WT* __restrict__ aw, int64_t n,
                                                   const uint32_t* __restrict__ bits, int64_t wpr, int64_t B,
                                                   double* __restrict__ q, int stage) {
    extern __shared__ uint32_t sbits[];
    __shared__ AT part[4];
    const int64_t b = blockIdx.x;
    const uint32_t* row = bits + b * wpr;
    const int64_t nw = (n + 31) >> 5;
    const uint32_t* xb = row;
    if (stage) {
        for (int64_t t = threadIdx.x; t < nw; t += blockDim.x) sbits[t] = row[t];
        __syncthreads();
        xb = sbits;
    }
    AT acc = 0;
    for (int64_t u = threadIdx.x; u < n; u += blockDim.x) {
        AT r = 0;
        for (int32_t e = ap[u]; e < ap[u + 1]; ++e) {
            const int32_t v = ai[e];
            const bool xv = (xb[v >> 5] >> (v & 31)) & 1u;
            r += xv ? -(AT)aw[e] : (AT)aw[e];
        }
        const bool xu = (xb[u >> 5] >> (u & 31)) & 1u;
        acc += xu ? -r : r;
    }
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) {
        if (sizeof(AT) == 8) {
            long long lv;
            __builtin_memcpy(&lv, &acc, 8);
            int lo = __shfl_xor((int)(lv & 0xffffffff), m, 64), hi = __shfl_xor((int)(lv >> 32), m, 64);
            long long ov = ((long long)hi << 32) | (unsigned)lo;
            AT o;
            __builtin_memcpy(&o, &ov, 8);
            acc += o;
        }
    }
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) q[b] = (double)(part[0] + part[1] + part[2] + part[3]);
}

size_t batch_scratch_bytes(const Graph& g, int64_t B) {
    if (g.W == nullptr) return 8;
    if (g.kind == PQI_WEIGHTS_I8) return (size_t)B * 8;
    int64_t nblk = (g.n + 63) / 64;
    if (B <= kFewMax) {
        int64_t nrb, ncs, cp;
        mfma_grid(g, &nrb, &ncs, &cp);
        nblk = std::max(nblk, nrb * ncs);
    }
    return (size_t)B * (size_t)nblk * 8;
}

// Waves per wide tile: 8 (two waves per SIMD; measured faster on MI355X),
// or 4 with PQI_WIDE_WAVES=4.
static int wide_waves() {
    static const int v = [] {
        const char* e = getenv("PQI_WIDE_WAVES");
        return e && atoi(e) == 4 ? 4 : 8;
    }();
    return v;
}

void launch_batch_i8_raw(const Graph& g, const uint32_t* bits, int64_t wpr, int64_t B, unsigned long long* qi,
                         hipStream_t s, const int32_t* gate) {
    if (B >= 512 && g.max_strength >= 0 && 256 * g.max_strength < (int64_t(1) << 31)) {
        const int64_t MB = (B + WIDE_BM - 1) / WIDE_BM, JB = (g.n_pad + WIDE_BN - 1) / WIDE_BN;
        // split K until there is about one work group per CU (each split >= 512 deep)
        int64_t KS = 1;
        const int64_t tiles = ((JB + 7) / 8) * 8 * MB, cus = g.ctx ? std::max(1, g.ctx->num_cus) : 256;
        while (tiles * KS * 2 <= cus && g.n_pad % (128 * KS * 2) == 0 && g.n_pad / (KS * 2) >= 512) KS *= 2;
        static const bool big =
            hipFuncSetAttribute((const void*)k_batch_i8_wide<128, 4>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)wide_lds_bytes<128>()) == hipSuccess &&
            hipFuncSetAttribute((const void*)k_batch_i8_wide<128, 8>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)wide_lds_bytes<128>()) == hipSuccess;
        const int nw = wide_waves();
        const dim3 grid((unsigned)(tiles * KS));
        if (big && nw == 8)
            hipLaunchKernelGGL((k_batch_i8_wide<128, 8>), grid, dim3(512), wide_lds_bytes<128>(), s,
                               (const int8_t*)g.W, g.n_pad, g.n_pad, bits, wpr, B, qi, MB, JB, KS, gate);
        else if (big)
            hipLaunchKernelGGL((k_batch_i8_wide<128, 4>), grid, dim3(256), wide_lds_bytes<128>(), s,
                               (const int8_t*)g.W, g.n_pad, g.n_pad, bits, wpr, B, qi, MB, JB, KS, gate);
        else if (nw == 8)
            hipLaunchKernelGGL((k_batch_i8_wide<64, 8>), grid, dim3(512), wide_lds_bytes<64>(), s,
                               (const int8_t*)g.W, g.n_pad, g.n_pad, bits, wpr, B, qi, MB, JB, KS, gate);
        else
            hipLaunchKernelGGL((k_batch_i8_wide<64, 4>), grid, dim3(256), wide_lds_bytes<64>(), s,
                               (const int8_t*)g.W, g.n_pad, g.n_pad, bits, wpr, B, qi, MB, JB, KS, gate);
    } else {
        dim3 grid((unsigned)(g.n_pad / BN), (unsigned)((B + BM - 1) / BM));
        hipLaunchKernelGGL(k_batch_i8, grid, dim3(256), 0, s, (const int8_t*)g.W, g.n_pad, g.n_pad, bits, wpr, B, qi,
                           gate);
    }
}

void launch_batch_quadratic(const Graph& g, const uint32_t* bits, int64_t wpr, int64_t B, double* q, void* scratch,
                            hipStream_t s) {
    if (B <= 0) return;
    if (g.W == nullptr) {
        const size_t lds = (size_t)((g.n + 31) >> 5) * 4;
        const int stage = lds <= 48 * 1024 ? 1 : 0;
        const size_t dyn = stage ? lds : 0;
        if (g.kind == PQI_WEIGHTS_I8)
            hipLaunchKernelGGL((k_batch_csr<int8_t, int64_t>), dim3((unsigned)B), dim3(256), dyn, s, g.adj_ptr,
                               g.adj_idx, (const int8_t*)g.adj_w, g.n, bits, wpr, B, q, stage);
        else if (g.kind == PQI_WEIGHTS_F32)
            hipLaunchKernelGGL((k_batch_csr<float, double>), dim3((unsigned)B), dim3(256), dyn, s, g.adj_ptr,
                               g.adj_idx, (const float*)g.adj_w, g.n, bits, wpr, B, q, stage);
        else
            hipLaunchKernelGGL((k_batch_csr<double, double>), dim3((unsigned)B), dim3(256), dyn, s, g.adj_ptr,
                               g.adj_idx, (const double*)g.adj_w, g.n, bits, wpr, B, q, stage);
        return;
    }
    if (g.kind == PQI_WEIGHTS_I8) {
        unsigned long long* qi = (unsigned long long*)scratch;
        hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, s, qi, B);
        launch_batch_i8_raw(g, bits, wpr, B, qi, s);
        hipLaunchKernelGGL(k_i64_to_f64, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, s, qi, B, q);
    } else if (B <= kFewMax) {
        int64_t nrb, ncs, cp;
        mfma_grid(g, &nrb, &ncs, &cp);
        double* part = (double*)scratch;
        const dim3 grid((unsigned)((nrb * ncs + 3) / 4));
        if (g.kind == PQI_WEIGHTS_F32)
            hipLaunchKernelGGL((k_rescore_mfma<float>), grid, dim3(256), 0, s, (const float*)g.W, g.n_pad, nrb, bits,
                               wpr, (int)B, ncs, cp, part);
        else
            hipLaunchKernelGGL((k_rescore_mfma<double>), grid, dim3(256), 0, s, (const double*)g.W, g.n_pad, nrb,
                               bits, wpr, (int)B, ncs, cp, part);
        hipLaunchKernelGGL(k_sum_parts, dim3((unsigned)B), dim3(256), 0, s, part, nrb * ncs, q);
    } else {
        const int64_t nblk = (g.n + 63) / 64;
        double* part = (double*)scratch;
        dim3 grid((unsigned)nblk, (unsigned)((B + 63) / 64));
        if (g.kind == PQI_WEIGHTS_F32)
            hipLaunchKernelGGL((k_batch_real<float>), grid, dim3(256), 0, s, (const float*)g.W, g.n_pad, g.n, bits,
                               wpr, B, part, nblk);
        else
            hipLaunchKernelGGL((k_batch_real<double>), grid, dim3(256), 0, s, (const double*)g.W, g.n_pad, g.n, bits,
                               wpr, B, part, nblk);
        hipLaunchKernelGGL(k_batch_sum, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, s, part, nblk, B, q);
    }
}

}  // namespace pqi
