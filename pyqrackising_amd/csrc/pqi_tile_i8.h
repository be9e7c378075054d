// The int8 MFMA tile shared by the single-instance and the grouped (batched)
// candidate-energy kernels.
//
// One work group (256 threads, four waves in 2 x 2) computes, for candidates
// m0..m0+127 and columns j0..j0+127, P = S W on v_mfma_i32_32x32x32_i8 (each
// wave 64 x 64 as 2 x 2 MFMA blocks) and adds sum_j S_mj P_mj of its columns to
// q[m] with one int64 atomic per row and wave column (exact, order-free).
// Candidate bits (LSB-first uint32 words) are unpacked to +-1 int8 in LDS.
// BK is the contraction depth per LDS stage (64 or 128).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pqi {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int TILE_BM = 128;
constexpr int TILE_BN = 128;

// 4 bits -> 4 bytes holding 0/1.
__device__ __forceinline__ uint32_t spread4(uint32_t nib) { return (nib * 0x00204081u) & 0x01010101u; }

// 0/1 bytes -> +-1 int8 bytes (x = 1 -> s = -1).
__device__ __forceinline__ uint32_t to_pm1(uint32_t bytes01) { return (bytes01 * 0xFEu) | 0x01010101u; }

template <int BK>
__device__ __forceinline__ void tile_i8(const int8_t* __restrict__ W, int64_t ld, int64_t K,
                                        const uint32_t* __restrict__ bits, int64_t wpr, int64_t B, int64_t m0,
                                        int64_t j0, unsigned long long* __restrict__ q, int8_t* As, int8_t* Bs,
                                        uint32_t* Sb) {
    constexpr int ROW = BK + 16;          // padded LDS row (bytes)
    constexpr int AW = BK / 64;           // candidate words per loader thread (2 threads per row)
    constexpr int BQ = BK / 32;           // uint4 of W per loader thread
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm = wid >> 1, wn = wid & 1;
    const int lr = tid >> 1, lh = tid & 1;  // loader row / half

    v16i acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[a][b][r] = 0;

    const int64_t am = m0 + lr;
    for (int64_t k0 = 0; k0 < K; k0 += BK) {
#pragma unroll
        for (int w = 0; w < AW; ++w) {
            const uint32_t word = (am < B) ? bits[am * wpr + (k0 >> 5) + lh * AW + w] : 0u;
            uint32_t o[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) o[t] = to_pm1(spread4((word >> (4 * t)) & 0xFu));
            uint4* dst = (uint4*)&As[lr * ROW + (lh * AW + w) * 32];
            dst[0] = make_uint4(o[0], o[1], o[2], o[3]);
            dst[1] = make_uint4(o[4], o[5], o[6], o[7]);
        }
        {
            const uint4* src = (const uint4*)(W + (j0 + lr) * ld + k0 + lh * (BK / 2));
            uint4 v[BQ];
#pragma unroll
            for (int t = 0; t < BQ; ++t) v[t] = src[t];
            uint4* dst = (uint4*)&Bs[lr * ROW + lh * (BK / 2)];
#pragma unroll
            for (int t = 0; t < BQ; ++t) dst[t] = v[t];
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; kk += 32) {
            v4i av[2], bv[2];
#pragma unroll
            for (int mi = 0; mi < 2; ++mi)
                av[mi] = *(const v4i*)&As[(wm * 64 + mi * 32 + (lane & 31)) * ROW + kk + 16 * (lane >> 5)];
#pragma unroll
            for (int nj = 0; nj < 2; ++nj)
                bv[nj] = *(const v4i*)&Bs[(wn * 64 + nj * 32 + (lane & 31)) * ROW + kk + 16 * (lane >> 5)];
#pragma unroll
            for (int mi = 0; mi < 2; ++mi)
#pragma unroll
                for (int nj = 0; nj < 2; ++nj)
                    acc[mi][nj] = __builtin_amdgcn_mfma_i32_32x32x32_i8(av[mi], bv[nj], acc[mi][nj], 0, 0, 0);
        }
        __syncthreads();
    }

    // candidate bits of this block's output columns
    for (int t = tid; t < TILE_BM * (TILE_BN / 32); t += 256) {
        const int r = t >> 2, c = t & 3;
        const int64_t m = m0 + r;
        Sb[t] = (m < B) ? bits[m * wpr + (j0 >> 5) + c] : 0u;
    }
    __syncthreads();

    // 32x32 accumulator layout: lane l, register r -> column l%32,
    // row 8*(r/4) + 4*(l/32) + r%4.
    int vals[32];
#pragma unroll
    for (int i = 0; i < 32; ++i) vals[i] = 0;
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int nj = 0; nj < 2; ++nj) {
            const int col = wn * 64 + nj * 32 + (lane & 31);
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = wm * 64 + mi * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                const uint32_t bit = (Sb[row * 4 + (col >> 5)] >> (col & 31)) & 1u;
                const int p = acc[mi][nj][r];
                vals[mi * 16 + r] += bit ? -p : p;
            }
        }
    // reduce-scatter over the 32 lanes that share rows
#pragma unroll
    for (int t = 0, half = 16; t < 5; ++t, half >>= 1) {
        const int up = (lane >> t) & 1;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            if (i < half) {
                const int send = up ? vals[i] : vals[i + half];
                const int keep = up ? vals[i + half] : vals[i];
                vals[i] = keep + __shfl_xor(send, 1 << t, 64);
            }
        }
    }
    const int qi = (lane & 1) * 16 + ((lane >> 1) & 1) * 8 + ((lane >> 2) & 1) * 4 + ((lane >> 3) & 1) * 2 +
                   ((lane >> 4) & 1);
    const int mi = qi >> 4, r = qi & 15;
    const int row = wm * 64 + mi * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    const int64_t m = m0 + row;
    if (m < B) atomicAdd(&q[m], (unsigned long long)(long long)vals[0]);
    __syncthreads();  // LDS is reused by the caller's next tile
}

// ---------------------------------------------------------------------------
// Wide tile: 256 candidates x 256 columns per work group.  Wave w owns rows
// 64w..64w+63 and all 256 columns (2 x 8 MFMA blocks, 256 accumulators): each
// unpacked +-1 fragment feeds 8 MFMAs, and W crosses L2 once per 256
// candidates.  A wave's candidate rows are its own: each lane loads the packed
// bit words of its rows into registers one chunk ahead and unpacks its
// fragments there.  W, shared by the four waves, goes through LDS in two
// stages with coalesced loads: while the waves run chunk c from one stage,
// each thread stores chunk c+1 (loaded one chunk earlier) into the other and
// loads chunk c+2; one barrier per chunk.  Columns past n_pad (the last tile
// of a row of W) read a valid row and are dropped.  The contraction may be
// split over k (the result is linear in it); each split adds its own part.
// LDS: 2 x 256 x (BK+16) bytes of W, plus 256 x 8 words for the epilogue.
// ---------------------------------------------------------------------------
constexpr int WIDE_BM = 256;
constexpr int WIDE_BN = 256;

template <int BK>
constexpr size_t wide_lds_bytes() {
    return 2 * (size_t)WIDE_BN * (BK + 16) + (size_t)WIDE_BM * 8 * 4;
}

// 16 candidate bits (k ascending) -> one MFMA fragment of 16 +-1 int8 bytes
__device__ __forceinline__ v4i unpack16(uint32_t half) {
    v4i a;
    a[0] = (int)to_pm1(spread4(half & 0xFu));
    a[1] = (int)to_pm1(spread4((half >> 4) & 0xFu));
    a[2] = (int)to_pm1(spread4((half >> 8) & 0xFu));
    a[3] = (int)to_pm1(spread4((half >> 12) & 0xFu));
    return a;
}

// one chunk of BK: NCB column blocks starting at cb0, both row blocks
template <int BK, int NCB>
__device__ __forceinline__ void wide_compute(v16i (&acc)[2][NCB], const uint32_t (&aw)[2][BK / 32], const int8_t* Ws,
                                             int lc, int hi, int cb0) {
    constexpr int ROW = BK + 16;
#pragma unroll
    for (int t = 0; t < BK / 32; ++t) {
        v4i av[2];
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) av[rb] = unpack16(hi ? (aw[rb][t] >> 16) : (aw[rb][t] & 0xFFFFu));
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) {
            const v4i bv = *(const v4i*)&Ws[((cb0 + cb) * 32 + lc) * ROW + 32 * t + 16 * hi];
#pragma unroll
            for (int rb = 0; rb < 2; ++rb)
                acc[rb][cb] = __builtin_amdgcn_mfma_i32_32x32x32_i8(av[rb], bv, acc[rb][cb], 0, 0, 0);
        }
    }
}

template <int BK>
__device__ __forceinline__ void wide_load_a(uint32_t (&aw)[2][BK / 32], const uint32_t* const (&arow)[2], int64_t k0) {
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int t = 0; t < BK / 32; ++t) aw[rb][t] = arow[rb][(k0 >> 5) + t];
}

// W loader, coalesced: BK/16 consecutive lanes read the BK bytes of one W row,
// so a load instruction covers NT / (BK/16) rows; pass t takes rows
// t * NT/(BK/16) + tid / (BK/16).  wbase[t] = W row of pass t + piece offset.
template <int NP>
__device__ __forceinline__ void wide_load_w(v4i (&wv)[NP], const int8_t* const (&wbase)[NP], int64_t k0) {
#pragma unroll
    for (int t = 0; t < NP; ++t) wv[t] = *(const v4i*)(wbase[t] + k0);
}

template <int NP>
__device__ __forceinline__ void wide_store_w(const v4i (&wv)[NP], int8_t* stage, const int (&woff)[NP]) {
#pragma unroll
    for (int t = 0; t < NP; ++t) *(v4i*)(stage + woff[t]) = wv[t];
}

// Contraction over k in [kb, ke) (a split of K = ld; multiples of BK).
// NW = 4: wave w owns candidate rows [64w, 64w+64) and all 8 column blocks
// (256 accumulator registers per lane, one wave per SIMD).  NW = 8: wave w
// owns rows [64(w mod 4), +64) and column blocks [4(w/4), +4) (128 accumulator
// registers, two waves per SIMD, so one wave's MFMAs cover the other's loads).
template <int BK, int NW = 4>
__device__ __forceinline__ void tile_i8_wide(const int8_t* __restrict__ W, int64_t ld, int64_t kb, int64_t ke,
                                             const uint32_t* __restrict__ bits,
                                             int64_t wpr, int64_t B, int64_t m0, int64_t j0,
                                             unsigned long long* __restrict__ q, unsigned char* lds) {
    constexpr int ROW = BK + 16;
    constexpr int STAGE = WIDE_BN * ROW;
    constexpr int NT = 64 * NW;         // threads
    constexpr int NCB = 32 / NW;        // column blocks per wave
    static_assert(BK % 32 == 0, "BK: a multiple of 32");
    static_assert(NW == 4 || NW == 8, "NW: 4 or 8 waves");
    int8_t* stage0 = (int8_t*)lds;
    int8_t* stage1 = stage0 + STAGE;
    uint32_t* Sb = (uint32_t*)(lds + 2 * STAGE);
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int hi = lane >> 5, lc = lane & 31;
    const int rg = wid & 3, cb0 = (wid >> 2) * NCB;
    // this lane's candidate rows (one per row block); rows past B read row 0
    // and are dropped in the epilogue, so the loads need no branch
    const uint32_t* arow[2];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
        const int64_t m = m0 + rg * 64 + rb * 32 + lc;
        arow[rb] = bits + (m < B ? m : 0) * wpr;
    }
    constexpr int LPR = BK / 16;        // lanes per W row in the loader
    constexpr int RPP = NT / LPR;       // rows per loader pass
    constexpr int NP = WIDE_BN / RPP;   // loader passes per chunk
    const int8_t* wbase[NP];
    int woff[NP];
#pragma unroll
    for (int t = 0; t < NP; ++t) {
        const int r = t * RPP + tid / LPR, piece = tid % LPR;
        const int64_t jrow = j0 + r < ld ? j0 + r : ld - 1;  // rows past n_pad: a valid row, dropped later
        wbase[t] = W + jrow * ld + 16 * piece;
        woff[t] = r * ROW + 16 * piece;
    }

    v16i acc[2][NCB];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < NCB; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[a][b][r] = 0;

    const int64_t nc = (ke - kb) / BK;
    auto chunk = [nc, kb](int64_t c) { return kb + (c < nc ? c : nc - 1) * BK; };  // past the end: the last
    uint32_t a0[2][BK / 32], a1[2][BK / 32];
    v4i w0[NP], w1[NP];
    wide_load_a<BK>(a0, arow, kb);
    wide_load_w<NP>(w0, wbase, kb);
    wide_load_w<NP>(w1, wbase, chunk(1));
    wide_store_w<NP>(w0, stage0, woff);
    wide_load_w<NP>(w0, wbase, chunk(2));
    __syncthreads();
    int64_t c = 0;
    for (; c + 1 < nc; c += 2) {
        wide_load_a<BK>(a1, arow, chunk(c + 1));
        wide_compute<BK, NCB>(acc, a0, stage0, lc, hi, cb0);
        wide_store_w<NP>(w1, stage1, woff);
        wide_load_w<NP>(w1, wbase, chunk(c + 3));
        __syncthreads();
        wide_load_a<BK>(a0, arow, chunk(c + 2));
        wide_compute<BK, NCB>(acc, a1, stage1, lc, hi, cb0);
        wide_store_w<NP>(w0, stage0, woff);
        wide_load_w<NP>(w0, wbase, chunk(c + 4));
        __syncthreads();
    }
    if (c < nc) wide_compute<BK, NCB>(acc, a0, stage0, lc, hi, cb0);

    // candidate bits of the tile's 256 output columns (words past n_pad: 0)
    const int64_t cwords = (ld - j0) >> 5;
    for (int t = tid; t < WIDE_BM * 8; t += NT) {
        const int r = t >> 3, cc = t & 7;
        const int64_t m = m0 + r;
        Sb[t] = (m < B && cc < cwords) ? bits[m * wpr + (j0 >> 5) + cc] : 0u;
    }
    __syncthreads();
    // 32x32 accumulator layout: lane l, register r -> column l%32,
    // row 8*(r/4) + 4*(l/32) + r%4 of the block.  P = c_j - 2 (X W).
    // column blocks past n_pad are dropped (uniform per block)
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb)
        if (j0 + (cb0 + cb) * 32 >= ld)
#pragma unroll
            for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[rb][cb][r] = 0;
    // per element: m = -(candidate bit) by one signed bit-field extract, and
    // s * P = (P + m) ^ m
    int vals[32];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = rg * 64 + rb * 32 + 8 * (r >> 2) + 4 * hi + (r & 3);
            const v4i w03 = *(const v4i*)&Sb[row * 8];
            const v4i w47 = *(const v4i*)&Sb[row * 8 + 4];
            int v = 0;
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb) {
                // this block's word of the row: block cb0 + cb (cb0 is uniform per wave)
                const int word = NW == 4 ? (cb < 4 ? w03[cb] : w47[cb - 4]) : (cb0 == 0 ? w03[cb] : w47[cb]);
                const int m = __builtin_amdgcn_sbfe(word, (unsigned)lc, 1u);
                v += (acc[rb][cb][r] + m) ^ m;
            }
            vals[rb * 16 + r] = v;
        }
    // reduce-scatter over the 32 lanes that share rows: 32 -> 1 value
#pragma unroll
    for (int t = 0, half = 16; t < 5; ++t, half >>= 1) {
        const int up = (lane >> t) & 1;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            if (i < half) {
                const int send = up ? vals[i] : vals[i + half];
                const int keep = up ? vals[i + half] : vals[i];
                vals[i] = keep + __shfl_xor(send, 1 << t, 64);
            }
        }
    }
    const int qi = (lane & 1) * 16 + ((lane >> 1) & 1) * 8 + ((lane >> 2) & 1) * 4 + ((lane >> 3) & 1) * 2 +
                   ((lane >> 4) & 1);
    const int rb = qi >> 4, r = qi & 15;
    const int64_t m = m0 + rg * 64 + rb * 32 + 8 * (r >> 2) + 4 * hi + (r & 3);
    if (m < B) atomicAdd(&q[m], (unsigned long long)(long long)vals[0]);
    __syncthreads();  // LDS is reused by the caller's next tile
}

}  // namespace pqi
