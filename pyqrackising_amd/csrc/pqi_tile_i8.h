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

}  // namespace pqi
