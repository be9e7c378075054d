// Flip-proposal network: a 3-layer message-passing net over the nodes of a
// batch of sparse graphs, with hand-written forward, backward and Adam kernels.
//
//   input  X0[v]  = (Delta_v / strength_v, degree_v / 16, s_v, min(age_v, 64) / 64)
//   layer l = 1..3:  M_l = A H_{l-1}   (A: weighted adjacency, block diagonal over graphs)
//                    Z_l = [H_{l-1} | M_l] Theta_l + b_l,   H_l = relu(Z_l)
//   output score_v = H_3 . theta_o + b_o   (logit that spin v should flip)
//   loss   mean over nodes of softplus(score) - y score
//
// Hidden width 64.  Layer transforms run either in FP32 (reference mode, used
// by the gradient checks) or on v_mfma_f32_32x32x16_bf16 with BF16 operands
// and FP32 accumulation (the default).  Backward is FP32; weight-gradient
// reductions over nodes are split in fixed-size row chunks and summed in a
// fixed order, so results are deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "pqi_internal.h"

namespace pqi {

constexpr int HID = 64;
constexpr int F0 = 4;
constexpr int CHUNK = 1024;  // rows per weight-gradient partial

// parameter offsets (floats)
constexpr int64_t OFF_T1 = 0;                        // (2 F0) x HID
constexpr int64_t OFF_B1 = OFF_T1 + 2 * F0 * HID;    // HID
constexpr int64_t OFF_T2 = OFF_B1 + HID;             // (2 HID) x HID
constexpr int64_t OFF_B2 = OFF_T2 + 2 * HID * HID;
constexpr int64_t OFF_T3 = OFF_B2 + HID;
constexpr int64_t OFF_B3 = OFF_T3 + 2 * HID * HID;
constexpr int64_t OFF_TO = OFF_B3 + HID;             // HID
constexpr int64_t OFF_BO = OFF_TO + HID;             // 1
constexpr int64_t NPARAM = OFF_BO + 1;

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// M[v][f] = sum_e w_e X[idx_e][f]; one wave per row, lanes over features.
__global__ __launch_bounds__(256) void k_spmm(const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                                              const float* __restrict__ w, const float* __restrict__ X, int d,
                                              int64_t N, float* __restrict__ M, int accumulate) {
    const int64_t v = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int f = threadIdx.x & 63;
    if (v >= N || f >= d) return;
    float acc = 0.f;
    for (int32_t e = ptr[v]; e < ptr[v + 1]; ++e) acc += w[e] * X[(int64_t)idx[e] * d + f];
    if (accumulate) M[v * d + f] += acc;
    else M[v * d + f] = acc;
}

// Z = [H | M] Theta + b (FP32), H and M of width din; optional relu into Hout.
__global__ __launch_bounds__(256) void k_linear_f32(const float* __restrict__ H, const float* __restrict__ M, int din,
                                                    const float* __restrict__ Th, const float* __restrict__ b,
                                                    int64_t N, float* __restrict__ Z, float* __restrict__ Hout) {
    __shared__ float th[2 * HID * HID];
    const int K = 2 * din;
    for (int t = threadIdx.x; t < K * HID; t += blockDim.x) th[t] = Th[t];
    __syncthreads();
    const int64_t v = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int j = threadIdx.x & 63;
    if (v >= N) return;
    float acc = b[j];
    const float* hv = H + v * din;
    const float* mv = M + v * din;
    for (int k = 0; k < din; ++k) acc = fmaf(hv[k], th[k * HID + j], acc);
    for (int k = 0; k < din; ++k) acc = fmaf(mv[k], th[(din + k) * HID + j], acc);
    Z[v * HID + j] = acc;
    Hout[v * HID + j] = acc > 0.f ? acc : 0.f;
}

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// Z = [H | M] Theta + b on BF16 MFMA (din = HID = 64, K = 128): a work group of
// four waves computes 128 rows x 64 columns; each wave 32 rows as two 32x32
// blocks.  Operands are rounded to BF16 while staged into LDS.
__global__ __launch_bounds__(256) void k_linear_bf16(const float* __restrict__ H, const float* __restrict__ M,
                                                     const float* __restrict__ Th, const float* __restrict__ b,
                                                     int64_t N, float* __restrict__ Z, float* __restrict__ Hout) {
    constexpr int K = 2 * HID;
    constexpr int ROW = K * 2 + 16;  // bytes per LDS row
    __shared__ __align__(16) unsigned char As[128 * ROW];
    __shared__ __align__(16) unsigned char Bs[HID * ROW];  // Theta transposed: column j, K contiguous
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int64_t r0 = (int64_t)blockIdx.x * 128;
    for (int t = tid; t < K * HID; t += 256) {
        const int k = t / HID, j = t % HID;
        *(uint16_t*)&Bs[j * ROW + k * 2] = f2bf(Th[t]);
    }
    for (int t = tid; t < 128 * K; t += 256) {
        const int r = t / K, k = t % K;
        const int64_t v = r0 + r;
        float val = 0.f;
        if (v < N) val = k < HID ? H[v * HID + k] : M[v * HID + (k - HID)];
        *(uint16_t*)&As[r * ROW + k * 2] = f2bf(val);
    }
    __syncthreads();
    v16f acc[2];
#pragma unroll
    for (int nj = 0; nj < 2; ++nj)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[nj][r] = 0.f;
#pragma unroll
    for (int ks = 0; ks < K; ks += 16) {
        const v8bf a = *(const v8bf*)&As[(wid * 32 + (lane & 31)) * ROW + (ks + 8 * (lane >> 5)) * 2];
#pragma unroll
        for (int nj = 0; nj < 2; ++nj) {
            const v8bf bb = *(const v8bf*)&Bs[(nj * 32 + (lane & 31)) * ROW + (ks + 8 * (lane >> 5)) * 2];
            acc[nj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bb, acc[nj], 0, 0, 0);
        }
    }
#pragma unroll
    for (int nj = 0; nj < 2; ++nj)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = wid * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            const int col = nj * 32 + (lane & 31);
            const int64_t v = r0 + row;
            if (v < N) {
                const float z = acc[nj][r] + b[col];
                Z[v * HID + col] = z;
                Hout[v * HID + col] = z > 0.f ? z : 0.f;
            }
        }
}

// ---------------------------------------------------------------------------
// Inference forward in BF16 mode (the learned search and pqi_mpnn_forward):
// one kernel per layer, activations stored as BF16, nothing kept for a
// backward pass.  Layer 1 (4 inputs) runs one thread per node; layers 2 and 3
// stage, per 128-row tile, each row's own activation and its neighbour sum
// (accumulated in FP32 from the BF16 neighbour rows, 8 threads per row, one
// 16-byte load per neighbour each) into LDS as BF16, then run the same
// 32x32x16 BF16 MFMA tile as k_linear_bf16; bias and relu are applied in the
// epilogue, and the last layer folds the output score (FP32) instead of
// writing its activations.  Tiles are dealt so that each XCD walks a
// contiguous run of rows (block b runs on XCD b mod 8): the rows of one graph
// share an XCD, so their neighbours' activations are read from that XCD's L2.
// ---------------------------------------------------------------------------

// tile of block b when the 8 XCDs each take a contiguous run of the tiles
// (the grid is 8 ceil(ntiles / 8) blocks; tiles >= ntiles are idle)
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t ntiles) {
    const int64_t per = (ntiles + 7) / 8;
    return (b % 8) * per + b / 8;
}

__device__ __forceinline__ float bf2f(uint32_t h) { return __uint_as_float(h << 16); }

// Theta_2 and Theta_3 to BF16, transposed: Tb[l][j][k] (column j, K contiguous)
__global__ void k_theta_bf16(const float* __restrict__ P, uint16_t* __restrict__ Tb) {
    constexpr int KH = 2 * HID * HID;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * KH) return;
    const int l = t / KH, r = t % KH;
    const int k = r / HID, j = r % HID;
    Tb[l * KH + j * (2 * HID) + k] = f2bf(P[(l == 0 ? OFF_T2 : OFF_T3) + r]);
}

// layer 1: H1 = relu([X | A X] Theta_1 + b_1), one thread per node, BF16 out
__global__ __launch_bounds__(256) void k_infer_l1(const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                                                  const float* __restrict__ w, const float* __restrict__ X,
                                                  const float* __restrict__ P, int64_t N, int64_t ntiles,
                                                  uint16_t* __restrict__ H1) {
    const int64_t t = xcd_tile(blockIdx.x, ntiles);
    if (t >= ntiles) return;
    const int64_t v = t * 256 + threadIdx.x;
    if (v >= N) return;
    const float4 x = *(const float4*)(X + v * F0);
    float m0 = 0.f, m1 = 0.f, m2 = 0.f, m3 = 0.f;
    int32_t e = ptr[v];
    const int32_t e1 = ptr[v + 1];
    for (; e + 4 <= e1; e += 4) {
        int32_t iu[4];
        float wu[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            iu[u] = idx[e + u];
            wu[u] = w[e + u];
        }
        float4 xu[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) xu[u] = *(const float4*)(X + (int64_t)iu[u] * F0);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            m0 += wu[u] * xu[u].x;
            m1 += wu[u] * xu[u].y;
            m2 += wu[u] * xu[u].z;
            m3 += wu[u] * xu[u].w;
        }
    }
    for (; e < e1; ++e) {
        const float we = w[e];
        const float4 xu = *(const float4*)(X + (int64_t)idx[e] * F0);
        m0 += we * xu.x;
        m1 += we * xu.y;
        m2 += we * xu.z;
        m3 += we * xu.w;
    }
    const float* T1 = P + OFF_T1;
    const float* b1 = P + OFF_B1;
    uint4 out[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        uint32_t pk[4];
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            uint32_t two = 0;
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int j = q * 8 + h * 2 + u;
                float z = b1[j];
                z = fmaf(x.x, T1[0 * HID + j], z);
                z = fmaf(x.y, T1[1 * HID + j], z);
                z = fmaf(x.z, T1[2 * HID + j], z);
                z = fmaf(x.w, T1[3 * HID + j], z);
                z = fmaf(m0, T1[4 * HID + j], z);
                z = fmaf(m1, T1[5 * HID + j], z);
                z = fmaf(m2, T1[6 * HID + j], z);
                z = fmaf(m3, T1[7 * HID + j], z);
                two |= (uint32_t)f2bf(z > 0.f ? z : 0.f) << (16 * u);
            }
            pk[h] = two;
        }
        out[q] = make_uint4(pk[0], pk[1], pk[2], pk[3]);
    }
    uint4* dst = (uint4*)(H1 + v * HID);
#pragma unroll
    for (int q = 0; q < 8; ++q) dst[q] = out[q];
}

// layers 2 and 3: relu([H | A H] Theta + b) on BF16 MFMA; LAST: score = . to + bo
template <bool LAST>
__global__ __launch_bounds__(256) void k_infer_layer(const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                                                     const float* __restrict__ w, const uint16_t* __restrict__ Hin,
                                                     const uint16_t* __restrict__ Tb, const float* __restrict__ bias,
                                                     const float* __restrict__ P, int64_t N, int64_t ntiles,
                                                     uint16_t* __restrict__ Hout, float* __restrict__ score) {
    constexpr int K = 2 * HID;
    constexpr int ROW = K * 2 + 16;  // bytes per LDS row of the operand tiles
    __shared__ __align__(16) unsigned char As[128 * ROW];
    __shared__ __align__(16) unsigned char Bs[HID * ROW];
    const int64_t t = xcd_tile(blockIdx.x, ntiles);
    if (t >= ntiles) return;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int64_t r0 = t * 128;
    // Theta (BF16, column j with K contiguous): 64 rows of 256 bytes
    for (int c = tid; c < HID * (K / 8); c += 256) {
        const int j = c / (K / 8), q = c % (K / 8);
        *(uint4*)&Bs[j * ROW + q * 16] = *(const uint4*)(Tb + (int64_t)j * K + q * 8);
    }
    // own activations: columns 0 .. 63
    for (int c = tid; c < 128 * 8; c += 256) {
        const int r = c >> 3, q = c & 7;
        const int64_t v = r0 + r;
        uint4 val = make_uint4(0u, 0u, 0u, 0u);
        if (v < N) val = *(const uint4*)(Hin + v * HID + q * 8);
        *(uint4*)&As[r * ROW + q * 16] = val;
    }
    // neighbour sums: columns 64 .. 127, 8 threads per row, 8 features each
    for (int r = tid >> 3; r < 128; r += 32) {
        const int q = tid & 7;
        const int64_t v = r0 + r;
        float acc[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) acc[i] = 0.f;
        if (v < N) {
            // eight neighbours at a time: their indices, then their rows, in flight together
            int32_t e = ptr[v];
            const int32_t e1 = ptr[v + 1];
            for (; e + 8 <= e1; e += 8) {
                int32_t iu[8];
                float wu[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    iu[u] = idx[e + u];
                    wu[u] = w[e + u];
                }
                uint4 hv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) hv[u] = *(const uint4*)(Hin + (int64_t)iu[u] * HID + q * 8);
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const uint32_t hw[4] = {hv[u].x, hv[u].y, hv[u].z, hv[u].w};
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        acc[2 * i] += wu[u] * bf2f(hw[i] & 0xFFFFu);
                        acc[2 * i + 1] += wu[u] * bf2f(hw[i] >> 16);
                    }
                }
            }
            if (e + 4 <= e1) {
                int32_t iu[4];
                float wu[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    iu[u] = idx[e + u];
                    wu[u] = w[e + u];
                }
                uint4 hv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) hv[u] = *(const uint4*)(Hin + (int64_t)iu[u] * HID + q * 8);
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const uint32_t hw[4] = {hv[u].x, hv[u].y, hv[u].z, hv[u].w};
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        acc[2 * i] += wu[u] * bf2f(hw[i] & 0xFFFFu);
                        acc[2 * i + 1] += wu[u] * bf2f(hw[i] >> 16);
                    }
                }
                e += 4;
            }
            for (; e < e1; ++e) {
                const float we = w[e];
                const uint4 hv = *(const uint4*)(Hin + (int64_t)idx[e] * HID + q * 8);
                const uint32_t hw[4] = {hv.x, hv.y, hv.z, hv.w};
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    acc[2 * i] += we * bf2f(hw[i] & 0xFFFFu);
                    acc[2 * i + 1] += we * bf2f(hw[i] >> 16);
                }
            }
        }
        uint32_t pk[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) pk[i] = (uint32_t)f2bf(acc[2 * i]) | ((uint32_t)f2bf(acc[2 * i + 1]) << 16);
        *(uint4*)&As[r * ROW + 128 + q * 16] = make_uint4(pk[0], pk[1], pk[2], pk[3]);
    }
    __syncthreads();
    v16f acc[2];
#pragma unroll
    for (int nj = 0; nj < 2; ++nj)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[nj][r] = 0.f;
#pragma unroll
    for (int ks = 0; ks < K; ks += 16) {
        const v8bf a = *(const v8bf*)&As[(wid * 32 + (lane & 31)) * ROW + (ks + 8 * (lane >> 5)) * 2];
#pragma unroll
        for (int nj = 0; nj < 2; ++nj) {
            const v8bf bb = *(const v8bf*)&Bs[(nj * 32 + (lane & 31)) * ROW + (ks + 8 * (lane >> 5)) * 2];
            acc[nj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bb, acc[nj], 0, 0, 0);
        }
    }
    __syncthreads();  // the operand tile is reused for the output
    if constexpr (LAST) {
        float* Of = (float*)As;  // 128 x 64 FP32 activations
#pragma unroll
        for (int nj = 0; nj < 2; ++nj)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = wid * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                const int col = nj * 32 + (lane & 31);
                const float z = acc[nj][r] + bias[col];
                Of[row * HID + col] = z > 0.f ? z : 0.f;
            }
        __syncthreads();
        if (tid < 128 && r0 + tid < N) {
            float sc = 0.f;
            for (int j = 0; j < HID; ++j) sc = fmaf(Of[tid * HID + j], P[OFF_TO + j], sc);
            score[r0 + tid] = sc + P[OFF_BO];
        }
    } else {
        uint16_t* Ob = (uint16_t*)As;  // 128 x 64 BF16 activations
#pragma unroll
        for (int nj = 0; nj < 2; ++nj)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = wid * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                const int col = nj * 32 + (lane & 31);
                const float z = acc[nj][r] + bias[col];
                Ob[row * HID + col] = f2bf(z > 0.f ? z : 0.f);
            }
        __syncthreads();
        for (int c = tid; c < 128 * 8; c += 256) {
            const int r = c >> 3, q = c & 7;
            const int64_t v = r0 + r;
            if (v < N) *(uint4*)(Hout + v * HID + q * 8) = *(const uint4*)&Ob[r * HID + q * 8];
        }
    }
}

// score = H3 . to + bo
__global__ __launch_bounds__(256) void k_out(const float* __restrict__ H, const float* __restrict__ P, int64_t N,
                                             float* __restrict__ score) {
    const int64_t v = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int j = threadIdx.x & 63;
    if (v >= N) return;
    float t = H[v * HID + j] * P[OFF_TO + j];
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) t += __shfl_xor(t, m, 64);
    if (j == 0) score[v] = t + P[OFF_BO];
}

// loss terms and dscore = (sigmoid(score) - y) / N
__global__ void k_bce(const float* __restrict__ score, const float* __restrict__ y, int64_t N, float* __restrict__ ds,
                      double* __restrict__ lpart) {
    __shared__ double sh[256];
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double l = 0.0;
    if (v < N) {
        const float s = score[v];
        const float sp = s > 0.f ? s + log1pf(expf(-s)) : log1pf(expf(s));
        l = (double)(sp - y[v] * s);
        const float sig = 1.f / (1.f + expf(-s));
        ds[v] = (sig - y[v]) / (float)N;
    }
    sh[threadIdx.x] = l;
    __syncthreads();
    for (int m = 128; m > 0; m >>= 1) {
        if (threadIdx.x < m) sh[threadIdx.x] += sh[threadIdx.x + m];
        __syncthreads();
    }
    if (threadIdx.x == 0) lpart[blockIdx.x] = sh[0];
}

// dH3 = ds * to; and chunk partials of dto = sum_v ds_v H3_v, dbo = sum ds
__global__ __launch_bounds__(64) void k_out_back(const float* __restrict__ H, const float* __restrict__ ds,
                                                 const float* __restrict__ P, int64_t N, float* __restrict__ dH,
                                                 float* __restrict__ part) {
    const int j = threadIdx.x;
    const int64_t c0 = (int64_t)blockIdx.x * CHUNK;
    float g = 0.f, gb = 0.f;
    for (int64_t v = c0; v < std::min<int64_t>(N, c0 + CHUNK); ++v) {
        const float d = ds[v];
        dH[v * HID + j] = d * P[OFF_TO + j];
        g += d * H[v * HID + j];
        gb += d;
    }
    part[blockIdx.x * (HID + 1) + j] = g;
    if (j == 0) part[blockIdx.x * (HID + 1) + HID] = gb;
}

// dZ = dH .* (Z > 0), in place into dH
__global__ void k_relu_back(const float* __restrict__ Z, float* __restrict__ dH, int64_t count) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count && !(Z[t] > 0.f)) dH[t] = 0.f;
}

// chunk partials of dTheta = [H | M]^T dZ (K x HID) and db = sum dZ
__global__ __launch_bounds__(256) void k_wgrad(const float* __restrict__ H, const float* __restrict__ M, int din,
                                               const float* __restrict__ dZ, int64_t N, float* __restrict__ part) {
    const int K = 2 * din;
    const int64_t c0 = (int64_t)blockIdx.x * CHUNK;
    const int64_t c1 = std::min<int64_t>(N, c0 + CHUNK);
    float* out = part + (int64_t)blockIdx.x * (K * HID + HID);
    for (int t = threadIdx.x; t < K * HID + HID; t += blockDim.x) {
        float acc = 0.f;
        if (t < K * HID) {
            const int k = t / HID, j = t % HID;
            const float* src = k < din ? H + k : M + (k - din);
            for (int64_t v = c0; v < c1; ++v) acc += src[v * din] * dZ[v * HID + j];
        } else {
            const int j = t - K * HID;
            for (int64_t v = c0; v < c1; ++v) acc += dZ[v * HID + j];
        }
        out[t] = acc;
    }
}

// sum chunk partials in fixed order into grad[off .. off + len)
__global__ void k_reduce_parts(const float* __restrict__ part, int64_t nchunks, int64_t len, float* __restrict__ grad) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= len) return;
    float acc = 0.f;
    for (int64_t c = 0; c < nchunks; ++c) acc += part[c * len + t];
    grad[t] = acc;
}

// dH_prev = dZ Theta[0:din]^T, dM = dZ Theta[din:2din]^T
__global__ __launch_bounds__(256) void k_input_grad(const float* __restrict__ dZ, const float* __restrict__ Th, int din,
                                                    int64_t N, float* __restrict__ dHp, float* __restrict__ dM) {
    __shared__ float th[2 * HID * HID];
    const int K = 2 * din;
    for (int t = threadIdx.x; t < K * HID; t += blockDim.x) th[t] = Th[t];
    __syncthreads();
    const int64_t v = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int k = threadIdx.x & 63;
    if (v >= N || k >= din) return;
    const float* dz = dZ + v * HID;
    float a = 0.f, m = 0.f;
    for (int j = 0; j < HID; ++j) {
        a = fmaf(dz[j], th[k * HID + j], a);
        m = fmaf(dz[j], th[(din + k) * HID + j], m);
    }
    dHp[v * din + k] = a;
    dM[v * din + k] = m;
}

__global__ void k_adam(float* __restrict__ P, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, int64_t n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float gt = g[t];
    const float a = b1 * m1[t] + (1.f - b1) * gt;
    const float v = b2 * m2[t] + (1.f - b2) * gt * gt;
    m1[t] = a;
    m2[t] = v;
    P[t] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

}  // namespace pqi

// ---------------------------------------------------------------------------
// host object
// ---------------------------------------------------------------------------

struct pqi_mpnn {
    pqi_ctx* ctx = nullptr;
    float* P = nullptr;  // NPARAM
    float* G = nullptr;
    float* m1 = nullptr;
    float* m2 = nullptr;
    int64_t steps = 0;
    int precision = 1;  // 0 fp32 layers, 1 bf16 MFMA layers
    // bound batch
    int64_t N = 0, E = 0;
    int32_t* ptr = nullptr;
    int32_t* idx = nullptr;
    float* w = nullptr;
    std::vector<int64_t> node_off;
    // activations
    float *X0 = nullptr, *M1 = nullptr, *Z1 = nullptr, *H1 = nullptr, *M2 = nullptr, *Z2 = nullptr, *H2 = nullptr;
    float *M3 = nullptr, *Z3 = nullptr, *H3 = nullptr, *score = nullptr, *y = nullptr, *ds = nullptr;
    float *dA = nullptr, *dB = nullptr, *dC = nullptr, *part = nullptr;
    double* lpart = nullptr;
    // the BF16 copy of Theta_2/3 in `part` matches P (cleared when P changes
    // or a backward pass reuses `part`)
    bool theta_bf16 = false;
};

namespace pqi {

static void mpnn_free_batch(pqi_mpnn* m) {
    m->theta_bf16 = false;
    float** fs[] = {&m->X0, &m->M1, &m->Z1, &m->H1, &m->M2, &m->Z2, &m->H2, &m->M3, &m->Z3,
                    &m->H3, &m->score, &m->y, &m->ds, &m->dA, &m->dB, &m->dC, &m->part, &m->w};
    for (float** f : fs)
        if (*f) {
            (void)hipFree(*f);
            *f = nullptr;
        }
    if (m->ptr) (void)hipFree(m->ptr);
    if (m->idx) (void)hipFree(m->idx);
    if (m->lpart) (void)hipFree(m->lpart);
    m->ptr = m->idx = nullptr;
    m->lpart = nullptr;
    m->N = m->E = 0;
}

// splitmix64: deterministic parameter initialisation
static uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static void forward(pqi_mpnn* m, hipStream_t s) {
    const int64_t N = m->N;
    const unsigned g4 = (unsigned)((N + 3) / 4);
    hipLaunchKernelGGL(k_spmm, dim3(g4), dim3(256), 0, s, m->ptr, m->idx, m->w, m->X0, F0, N, m->M1, 0);
    hipLaunchKernelGGL(k_linear_f32, dim3(g4), dim3(256), 0, s, m->X0, m->M1, F0, m->P + OFF_T1, m->P + OFF_B1, N,
                       m->Z1, m->H1);
    float* Hs[3] = {m->H1, m->H2, m->H3};
    float* Ms[3] = {m->M1, m->M2, m->M3};
    float* Zs[3] = {m->Z1, m->Z2, m->Z3};
    const int64_t offT[3] = {OFF_T1, OFF_T2, OFF_T3}, offB[3] = {OFF_B1, OFF_B2, OFF_B3};
    for (int l = 1; l < 3; ++l) {
        hipLaunchKernelGGL(k_spmm, dim3(g4), dim3(256), 0, s, m->ptr, m->idx, m->w, Hs[l - 1], HID, N, Ms[l], 0);
        if (m->precision == 1)
            hipLaunchKernelGGL(k_linear_bf16, dim3((unsigned)((N + 127) / 128)), dim3(256), 0, s, Hs[l - 1], Ms[l],
                               m->P + offT[l], m->P + offB[l], N, Zs[l], Hs[l]);
        else
            hipLaunchKernelGGL(k_linear_f32, dim3(g4), dim3(256), 0, s, Hs[l - 1], Ms[l], HID, m->P + offT[l],
                               m->P + offB[l], N, Zs[l], Hs[l]);
    }
    hipLaunchKernelGGL(k_out, dim3(g4), dim3(256), 0, s, m->H3, m->P, N, m->score);
}

// Scores for the bound batch from m->X0 (BF16 mode: the fused inference
// kernels, BF16 activations in the first halves of H1 and H2, Theta_2/3 as
// BF16 in the gradient-partial buffer, which only a backward pass uses; FP32
// mode: the training forward).
static void forward_infer(pqi_mpnn* m, hipStream_t s) {
    if (m->precision != 1) {
        forward(m, s);
        return;
    }
    const int64_t N = m->N;
    uint16_t* Tb = (uint16_t*)m->part;
    if (!m->theta_bf16) {
        hipLaunchKernelGGL(k_theta_bf16, dim3((unsigned)((4 * HID * HID + 255) / 256)), dim3(256), 0, s, m->P, Tb);
        m->theta_bf16 = true;
    }
    const int64_t t1 = (N + 255) / 256, tl = (N + 127) / 128;
    hipLaunchKernelGGL(k_infer_l1, dim3((unsigned)(8 * ((t1 + 7) / 8))), dim3(256), 0, s, m->ptr, m->idx, m->w, m->X0,
                       m->P, N, t1, (uint16_t*)m->H1);
    hipLaunchKernelGGL(k_infer_layer<false>, dim3((unsigned)(8 * ((tl + 7) / 8))), dim3(256), 0, s, m->ptr, m->idx,
                       m->w, (const uint16_t*)m->H1, Tb, m->P + OFF_B2, m->P, N, tl, (uint16_t*)m->H2, nullptr);
    hipLaunchKernelGGL(k_infer_layer<true>, dim3((unsigned)(8 * ((tl + 7) / 8))), dim3(256), 0, s, m->ptr, m->idx,
                       m->w, (const uint16_t*)m->H2, Tb + 2 * HID * HID, m->P + OFF_B3, m->P, N, tl, nullptr,
                       m->score);
}

// loss and gradient into m->G; returns the loss through *loss (after sync)
static int backward(pqi_mpnn* m, hipStream_t s, double* loss) {
    m->theta_bf16 = false;
    const int64_t N = m->N;
    const unsigned g4 = (unsigned)((N + 3) / 4);
    const int64_t nb = (N + 255) / 256;
    const int64_t nchunks = (N + CHUNK - 1) / CHUNK;
    hipLaunchKernelGGL(k_bce, dim3((unsigned)nb), dim3(256), 0, s, m->score, m->y, N, m->ds, m->lpart);
    // output layer
    hipLaunchKernelGGL(k_out_back, dim3((unsigned)nchunks), dim3(64), 0, s, m->H3, m->ds, m->P, N, m->dA, m->part);
    hipLaunchKernelGGL(k_reduce_parts, dim3((unsigned)((HID + 1 + 255) / 256)), dim3(256), 0, s, m->part, nchunks,
                       (int64_t)(HID + 1), m->G + OFF_TO);
    float* Hs[4] = {m->X0, m->H1, m->H2, m->H3};
    float* Ms[3] = {m->M1, m->M2, m->M3};
    float* Zs[3] = {m->Z1, m->Z2, m->Z3};
    const int64_t offT[3] = {OFF_T1, OFF_T2, OFF_T3};
    float* dH = m->dA;  // gradient w.r.t. H_l, then dZ_l in place
    for (int l = 2; l >= 0; --l) {
        const int din = l == 0 ? F0 : HID;
        const int K = 2 * din;
        hipLaunchKernelGGL(k_relu_back, dim3((unsigned)((N * HID + 255) / 256)), dim3(256), 0, s, Zs[l], dH, N * HID);
        hipLaunchKernelGGL(k_wgrad, dim3((unsigned)nchunks), dim3(256), 0, s, Hs[l], Ms[l], din, dH, N, m->part);
        hipLaunchKernelGGL(k_reduce_parts, dim3((unsigned)((K * HID + HID + 255) / 256)), dim3(256), 0, s, m->part,
                           nchunks, (int64_t)(K * HID + HID), m->G + offT[l]);
        if (l == 0) break;
        float* dHp = dH == m->dA ? m->dB : m->dA;
        hipLaunchKernelGGL(k_input_grad, dim3(g4), dim3(256), 0, s, dH, m->P + offT[l], din, N, dHp, m->dC);
        // dH_{l-1} += A^T dM  (A symmetric)
        hipLaunchKernelGGL(k_spmm, dim3(g4), dim3(256), 0, s, m->ptr, m->idx, m->w, m->dC, din, N, dHp, 1);
        dH = dHp;
    }
    std::vector<double> lp(nb);
    if (hipMemcpyAsync(lp.data(), m->lpart, nb * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return fail(PQI_ERR_HIP, "MPNN backward failed");
    double t = 0.0;
    for (double v : lp) t += v;
    if (loss) *loss = t / (double)N;
    return PQI_OK;
}

static int upload_inputs(pqi_mpnn* m, const float* feats, const float* labels) {
    hipStream_t s = m->ctx->stream;
    if (feats && hipMemcpyAsync(m->X0, feats, m->N * F0 * 4, hipMemcpyHostToDevice, s) != hipSuccess)
        return fail(PQI_ERR_HIP, "uploading features failed");
    if (labels && hipMemcpyAsync(m->y, labels, m->N * 4, hipMemcpyHostToDevice, s) != hipSuccess)
        return fail(PQI_ERR_HIP, "uploading labels failed");
    return PQI_OK;
}

}  // namespace pqi

namespace pqi {
float* mpnn_features(pqi_mpnn* m) { return m->X0; }
const float* mpnn_scores(pqi_mpnn* m) { return m->score; }
int64_t mpnn_node_offset(pqi_mpnn* m, int64_t i) {
    return i < (int64_t)m->node_off.size() ? m->node_off[i] : -1;
}
int64_t mpnn_total_nodes(pqi_mpnn* m) { return m->N; }
void mpnn_forward_device(pqi_mpnn* m, hipStream_t s) { forward_infer(m, s); }
}  // namespace pqi

using namespace pqi;

extern "C" {

int pqi_mpnn_create(pqi_ctx* c, uint64_t seed, pqi_mpnn** out) {
    if (!c || !out) return fail(PQI_ERR_INVALID, "null argument");
    if (hipSetDevice(c->device) != hipSuccess) return fail(PQI_ERR_HIP, "hipSetDevice failed");
    pqi_mpnn* m = new pqi_mpnn();
    m->ctx = c;
    if (hipMalloc(&m->P, NPARAM * 4) != hipSuccess || hipMalloc(&m->G, NPARAM * 4) != hipSuccess ||
        hipMalloc(&m->m1, NPARAM * 4) != hipSuccess || hipMalloc(&m->m2, NPARAM * 4) != hipSuccess) {
        delete m;
        return fail(PQI_ERR_OOM, "MPNN parameter allocation failed");
    }
    // uniform(-a, a) with a = sqrt(6 / (fan_in + fan_out)); biases 0
    std::vector<float> p(NPARAM, 0.f);
    uint64_t st = seed;
    auto fill = [&](int64_t off, int fan_in, int fan_out) {
        const double a = std::sqrt(6.0 / (fan_in + fan_out));
        for (int64_t t = 0; t < (int64_t)fan_in * fan_out; ++t) {
            const double u = (double)(splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
            p[off + t] = (float)((2.0 * u - 1.0) * a);
        }
    };
    fill(OFF_T1, 2 * F0, HID);
    fill(OFF_T2, 2 * HID, HID);
    fill(OFF_T3, 2 * HID, HID);
    fill(OFF_TO, HID, 1);
    std::vector<float> zeros(NPARAM, 0.f);
    if (hipMemcpy(m->P, p.data(), NPARAM * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->m1, zeros.data(), NPARAM * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->m2, zeros.data(), NPARAM * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->G, zeros.data(), NPARAM * 4, hipMemcpyHostToDevice) != hipSuccess) {
        delete m;
        return fail(PQI_ERR_HIP, "MPNN initialisation failed");
    }
    *out = m;
    return PQI_OK;
}

int pqi_mpnn_free(pqi_mpnn* m) {
    if (!m) return PQI_OK;
    (void)hipSetDevice(m->ctx->device);
    (void)hipStreamSynchronize(m->ctx->stream);
    mpnn_free_batch(m);
    (void)hipFree(m->P);
    (void)hipFree(m->G);
    (void)hipFree(m->m1);
    (void)hipFree(m->m2);
    delete m;
    return PQI_OK;
}

int pqi_mpnn_num_params(void) { return (int)NPARAM; }

int pqi_mpnn_set_precision(pqi_mpnn* m, int precision) {
    if (!m || (precision != 0 && precision != 1)) return fail(PQI_ERR_INVALID, "precision must be 0 (fp32) or 1 (bf16)");
    m->precision = precision;
    return PQI_OK;
}

int pqi_mpnn_get_params(pqi_mpnn* m, float* out) {
    if (!m || !out) return fail(PQI_ERR_INVALID, "null argument");
    if (hipMemcpy(out, m->P, NPARAM * 4, hipMemcpyDeviceToHost) != hipSuccess) return fail(PQI_ERR_HIP, "copy failed");
    return PQI_OK;
}

int pqi_mpnn_set_params(pqi_mpnn* m, const float* in) {
    if (!m || !in) return fail(PQI_ERR_INVALID, "null argument");
    m->theta_bf16 = false;
    if (hipMemcpy(m->P, in, NPARAM * 4, hipMemcpyHostToDevice) != hipSuccess) return fail(PQI_ERR_HIP, "copy failed");
    std::vector<float> zeros(NPARAM, 0.f);
    (void)hipMemcpy(m->m1, zeros.data(), NPARAM * 4, hipMemcpyHostToDevice);
    (void)hipMemcpy(m->m2, zeros.data(), NPARAM * 4, hipMemcpyHostToDevice);
    m->steps = 0;
    return PQI_OK;
}

// Bind a batch of graphs: their symmetric adjacencies are concatenated into
// one block-diagonal CSR (fp32 weights).
int pqi_mpnn_bind(pqi_mpnn* m, pqi_graph* const* graphs, int64_t count, int64_t* total_nodes) {
    if (!m || !graphs || count < 1) return fail(PQI_ERR_INVALID, "bad argument");
    if (hipSetDevice(m->ctx->device) != hipSuccess) return fail(PQI_ERR_HIP, "hipSetDevice failed");
    mpnn_free_batch(m);
    int64_t N = 0, E = 0;
    for (int64_t i = 0; i < count; ++i) {
        if (!graphs[i] || !graphs[i]->adj_ptr) return fail(PQI_ERR_UNSUPPORTED, "the MPNN needs adjacency lists");
        N += graphs[i]->n;
        E += graphs[i]->nnz;
    }
    if (N >= INT32_MAX || E >= INT32_MAX) return fail(PQI_ERR_UNSUPPORTED, "batch too large");
    std::vector<int32_t> ptr(N + 1, 0), idx(E);
    std::vector<float> w(E);
    m->node_off.assign(count + 1, 0);
    int64_t v0 = 0, e0 = 0;
    for (int64_t i = 0; i < count; ++i) {
        const pqi_graph* g = graphs[i];
        std::vector<int32_t> gp(g->n + 1), gi(g->nnz);
        std::vector<double> gw(g->nnz);
        if (hipMemcpy(gp.data(), g->adj_ptr, (g->n + 1) * 4, hipMemcpyDeviceToHost) != hipSuccess ||
            (g->nnz && hipMemcpy(gi.data(), g->adj_idx, g->nnz * 4, hipMemcpyDeviceToHost) != hipSuccess))
            return fail(PQI_ERR_HIP, "reading adjacency failed");
        if (g->nnz) {
            if (g->kind == PQI_WEIGHTS_I8) {
                std::vector<int8_t> t(g->nnz);
                if (hipMemcpy(t.data(), g->adj_w, g->nnz, hipMemcpyDeviceToHost) != hipSuccess)
                    return fail(PQI_ERR_HIP, "reading weights failed");
                for (int64_t e = 0; e < g->nnz; ++e) gw[e] = t[e];
            } else if (g->kind == PQI_WEIGHTS_F32) {
                std::vector<float> t(g->nnz);
                if (hipMemcpy(t.data(), g->adj_w, g->nnz * 4, hipMemcpyDeviceToHost) != hipSuccess)
                    return fail(PQI_ERR_HIP, "reading weights failed");
                for (int64_t e = 0; e < g->nnz; ++e) gw[e] = t[e];
            } else if (hipMemcpy(gw.data(), g->adj_w, g->nnz * 8, hipMemcpyDeviceToHost) != hipSuccess) {
                return fail(PQI_ERR_HIP, "reading weights failed");
            }
        }
        for (int64_t v = 0; v < g->n; ++v) ptr[v0 + v + 1] = (int32_t)(e0 + gp[v + 1]);
        for (int64_t e = 0; e < g->nnz; ++e) {
            idx[e0 + e] = (int32_t)(v0 + gi[e]);
            w[e0 + e] = (float)gw[e];
        }
        m->node_off[i] = v0;
        v0 += g->n;
        e0 += g->nnz;
    }
    m->node_off[count] = v0;
    m->N = N;
    m->E = E;
    const int64_t nchunks = (N + CHUNK - 1) / CHUNK;
    const size_t NH = (size_t)N * HID * 4;
    bool ok = hipMalloc(&m->ptr, (N + 1) * 4) == hipSuccess && hipMalloc(&m->idx, std::max<int64_t>(1, E) * 4) == hipSuccess &&
              hipMalloc(&m->w, std::max<int64_t>(1, E) * 4) == hipSuccess && hipMalloc(&m->X0, N * F0 * 4) == hipSuccess &&
              hipMalloc(&m->M1, N * F0 * 4) == hipSuccess && hipMalloc(&m->Z1, NH) == hipSuccess &&
              hipMalloc(&m->H1, NH) == hipSuccess && hipMalloc(&m->M2, NH) == hipSuccess &&
              hipMalloc(&m->Z2, NH) == hipSuccess && hipMalloc(&m->H2, NH) == hipSuccess &&
              hipMalloc(&m->M3, NH) == hipSuccess && hipMalloc(&m->Z3, NH) == hipSuccess &&
              hipMalloc(&m->H3, NH) == hipSuccess && hipMalloc(&m->score, N * 4) == hipSuccess &&
              hipMalloc(&m->y, N * 4) == hipSuccess && hipMalloc(&m->ds, N * 4) == hipSuccess &&
              hipMalloc(&m->dA, NH) == hipSuccess && hipMalloc(&m->dB, NH) == hipSuccess &&
              hipMalloc(&m->dC, NH) == hipSuccess &&
              hipMalloc(&m->part, (size_t)nchunks * (2 * HID * HID + HID) * 4) == hipSuccess &&
              hipMalloc(&m->lpart, ((N + 255) / 256) * 8) == hipSuccess;
    if (!ok) {
        mpnn_free_batch(m);
        return fail(PQI_ERR_OOM, "MPNN batch allocation failed");
    }
    if (hipMemcpy(m->ptr, ptr.data(), (N + 1) * 4, hipMemcpyHostToDevice) != hipSuccess ||
        (E && hipMemcpy(m->idx, idx.data(), E * 4, hipMemcpyHostToDevice) != hipSuccess) ||
        (E && hipMemcpy(m->w, w.data(), E * 4, hipMemcpyHostToDevice) != hipSuccess)) {
        mpnn_free_batch(m);
        return fail(PQI_ERR_HIP, "uploading the MPNN batch failed");
    }
    if (total_nodes) *total_nodes = N;
    return PQI_OK;
}

// scores (N) for host features (N x 4)
int pqi_mpnn_forward(pqi_mpnn* m, const float* feats, float* scores) {
    if (!m || !feats || !scores || !m->N) return fail(PQI_ERR_INVALID, "bad argument (bind a batch first)");
    if (hipSetDevice(m->ctx->device) != hipSuccess) return fail(PQI_ERR_HIP, "hipSetDevice failed");
    int r = upload_inputs(m, feats, nullptr);
    if (r != PQI_OK) return r;
    forward_infer(m, m->ctx->stream);
    if (hipMemcpyAsync(scores, m->score, m->N * 4, hipMemcpyDeviceToHost, m->ctx->stream) != hipSuccess ||
        hipStreamSynchronize(m->ctx->stream) != hipSuccess)
        return fail(PQI_ERR_HIP, "MPNN forward failed");
    return PQI_OK;
}

// loss and gradient (NPARAM floats, optional) for host features and labels
int pqi_mpnn_loss_grad(pqi_mpnn* m, const float* feats, const float* labels, double* loss, float* grad) {
    if (!m || !feats || !labels || !m->N) return fail(PQI_ERR_INVALID, "bad argument (bind a batch first)");
    if (hipSetDevice(m->ctx->device) != hipSuccess) return fail(PQI_ERR_HIP, "hipSetDevice failed");
    int r = upload_inputs(m, feats, labels);
    if (r != PQI_OK) return r;
    forward(m, m->ctx->stream);
    r = backward(m, m->ctx->stream, loss);
    if (r != PQI_OK) return r;
    if (grad && hipMemcpy(grad, m->G, NPARAM * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(PQI_ERR_HIP, "copy failed");
    return PQI_OK;
}

// one Adam step (beta 0.9 / 0.999, eps 1e-8)
int pqi_mpnn_train_step(pqi_mpnn* m, const float* feats, const float* labels, float lr, double* loss) {
    int r = pqi_mpnn_loss_grad(m, feats, labels, loss, nullptr);
    if (r != PQI_OK) return r;
    m->theta_bf16 = false;  // P changes below
    m->steps++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)m->steps), c2 = 1.f - std::pow(b2, (float)m->steps);
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((NPARAM + 255) / 256)), dim3(256), 0, m->ctx->stream, m->P, m->G, m->m1,
                       m->m2, NPARAM, lr, b1, b2, 1e-8f, c1, c2);
    if (hipStreamSynchronize(m->ctx->stream) != hipSuccess) return fail(PQI_ERR_HIP, "Adam step failed");
    return PQI_OK;
}

}  // extern "C"
