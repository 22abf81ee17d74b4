// sp_mlp.hip -- the f(p) surrogate: a dense SiLU network trained and run
// with BF16 matrix-core (MFMA) GEMMs accumulating in FP32, AdamW in FP32.
//
// Layer l maps K_l inputs to N_l outputs: Z_l = X_l W_l + b_l, with W_l
// (K_l x N_l, row-major) and b_l in one FP32 master vector (W_0, b_0, W_1,
// b_1, ...) and a BF16 copy that the GEMMs read.  Hidden layers store Z_l
// (FP32) for the backward pass and X_{l+1} = bf16(silu(Z_l)); the last layer
// is linear.  Rounding points (all round-to-nearest-even to BF16): the
// network input, every weight, every hidden activation, the output gradient
// and every back-propagated pre-activation gradient; all products accumulate
// in FP32.
//
// GEMM (training): 128 x 128, 128 x 64 or 64 x 64 output tiles per 256-lane
// block, four waves of a quarter tile each, K in steps of 32 staged through LDS
// as k-contiguous BF16 rows with 16-byte loads, v_mfma_f32_32x32x16_bf16 on the
// wave's 32 x 32 tiles.  Inference of the usual shapes runs all layers in one
// fused kernel instead (mlp_infer_kernel).  Operand lane layout (checked
// on gfx950): lane l holds row/column l % 32, k = 8 (l / 32) .. + 7; result
// register j of lane l is row 8 (j / 4) + 4 (l / 32) + j % 4, column l % 32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <string>
#include <vector>

#include "sp_device.h"
#include "sp_internal.h"

namespace sp {
namespace mlp {

typedef float v16f __attribute__((ext_vector_type(16)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

constexpr int kKT = 32;
constexpr int kLdsPad = 8;   // row stride 40 halves = 80 B

// Round to nearest even (v_cvt_pk_bf16_f32 on gfx950).
__device__ __forceinline__ uint16_t bf16_rne(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ __forceinline__ float bf16_to_f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// Two values rounded as bf16_rne, lo in the low half (one v_cvt_pk_bf16_f32).
__device__ __forceinline__ uint32_t bf16x2_rne(float lo, float hi) {
    typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
    typedef float v2f __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((v2f){lo, hi}, v2bf));
}

// 1 / (1 + e^-z) from the hardware exp2 and reciprocal (about 1 ulp each,
// far below the BF16 rounding that follows); z -> -inf gives 0, z -> inf 1.
__device__ __forceinline__ float sigmoid(float z) {
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.44269504088896341f * z));
}

enum Epilogue : int {
    EPI_HIDDEN = 0,   // z = acc + b: store z (FP32) and bf16(silu(z))
    EPI_INFER = 4,    // as EPI_HIDDEN without z (no backward pass follows)
    EPI_LINEAR = 1,   // y = acc + b: store y (FP32)
    EPI_BACK = 2,     // g = acc * silu'(zprev): store bf16(g)
    EPI_PLAIN = 3,    // store acc (FP32)
};

struct GemmArgs {
    const uint16_t* A;
    const uint16_t* B;
    int M, N, K;
    int kchunk;              // K range of one blockIdx.z (split K); 0: all of K
    int lda, ldb;
    int transA, transB;      // A(m,k) = transA ? A[k lda + m] : A[m lda + k]; same for B(k,n)
    const float* bias;
    const float* zprev;      // EPI_BACK: pre-activations, M x N
    float* outF;
    uint16_t* outH;
};

// Stages one operand tile (R rows or columns x kKT k) into LDS as k-contiguous
// rows.  X(r, k) = kmajor ? X[r ld + k] : X[k ld + r].  Chunks of 8 elements
// along the contiguous axis move as one 16-byte load when they are in range
// and aligned; the rest element by element (zero outside [0, nr) x [kb, ke)).
template <int R>
__device__ __forceinline__ void stage(uint16_t (*S)[kKT + kLdsPad], const uint16_t* X, int ld,
                                      bool kmajor, int r0, int nr, int k0, int ke) {
    const int t = threadIdx.x;
    constexpr int kChunks = R * kKT / 8;
    const bool vec_ok = (ld % 8) == 0 && (reinterpret_cast<uintptr_t>(X) % 16) == 0;
#pragma unroll
    for (int c = t; c < kChunks; c += 256) {
        if (kmajor) {
            const int r = c / (kKT / 8), kc = (c % (kKT / 8)) * 8;
            const int gr = r0 + r, gk = k0 + kc;
            if (vec_ok && gr < nr && gk + 8 <= ke) {
                *reinterpret_cast<uint4*>(&S[r][kc]) =
                    *reinterpret_cast<const uint4*>(X + (long long)gr * ld + gk);
            } else {
#pragma unroll
                for (int i = 0; i < 8; ++i)
                    S[r][kc + i] = (gr < nr && gk + i < ke) ? X[(long long)gr * ld + gk + i] : (uint16_t)0;
            }
        } else {
            const int k = c / (R / 8), rc = (c % (R / 8)) * 8;
            const int gk = k0 + k, gr = r0 + rc;
            if (vec_ok && gk < ke && gr + 8 <= nr) {
                const uint4 v = *reinterpret_cast<const uint4*>(X + (long long)gk * ld + gr);
                const uint16_t* h = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
                for (int i = 0; i < 8; ++i) S[rc + i][k] = h[i];
            } else {
#pragma unroll
                for (int i = 0; i < 8; ++i)
                    S[rc + i][k] = (gk < ke && gr + i < nr) ? X[(long long)gk * ld + gr + i] : (uint16_t)0;
            }
        }
    }
}

// BM x BN output tile per 256-lane block; wave (wm, wn) owns BM/2 x BN/2 as
// BM/64 x BN/64 MFMA tiles of 32 x 32.
template <int EPI, int BN, int BM>
__global__ void __launch_bounds__(256) gemm_bf16_kernel(GemmArgs g) {
    constexpr int TM = BM / 64;   // 32-row MFMA tiles per wave
    constexpr int TN = BN / 64;   // 32-column MFMA tiles per wave
    __shared__ __align__(16) uint16_t As[BM][kKT + kLdsPad];
    __shared__ __align__(16) uint16_t Bs[BN][kKT + kLdsPad];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wm = w >> 1, wn = w & 1;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    // split K: block z covers [kb, ke) and writes its own partial (EPI_PLAIN)
    const int kb = g.kchunk ? (int)blockIdx.z * g.kchunk : 0;
    const int ke = g.kchunk ? min(g.K, kb + g.kchunk) : g.K;
    // the bias starts the accumulation (the fused inference kernel does the same)
    v16f acc[TM][TN];
#pragma unroll
    for (int j = 0; j < TN; ++j) {
        const int col = n0 + wn * (BN / 2) + j * 32 + (lane & 31);
        const float b0 = (EPI == EPI_HIDDEN || EPI == EPI_INFER || EPI == EPI_LINEAR) && col < g.N ? g.bias[col] : 0.0f;
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = b0;
    }
    for (int k0 = kb; k0 < ke; k0 += kKT) {
        stage<BM>(As, g.A, g.lda, !g.transA, m0, g.M, k0, ke);
        stage<BN>(Bs, g.B, g.ldb, g.transB != 0, n0, g.N, k0, ke);
        __syncthreads();
#pragma unroll
        for (int kc = 0; kc < kKT; kc += 16) {
            v8bf a[TM], b[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i)
                a[i] = *reinterpret_cast<const v8bf*>(&As[wm * (BM / 2) + i * 32 + (lane & 31)][kc + 8 * (lane >> 5)]);
#pragma unroll
            for (int j = 0; j < TN; ++j)
                b[j] = *reinterpret_cast<const v8bf*>(&Bs[wn * (BN / 2) + j * 32 + (lane & 31)][kc + 8 * (lane >> 5)]);
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < TM; ++i) {
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const int col = n0 + wn * (BN / 2) + j * 32 + (lane & 31);
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = m0 + wm * (BM / 2) + i * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
                if (row >= g.M || col >= g.N) continue;
                const long long o = (long long)row * g.N + col;
                const float v = acc[i][j][r];
                if (EPI == EPI_HIDDEN || EPI == EPI_INFER) {
                    if (EPI == EPI_HIDDEN) g.outF[o] = v;
                    g.outH[o] = bf16_rne(v * sigmoid(v));
                } else if (EPI == EPI_LINEAR) {
                    g.outF[o] = v;
                } else if (EPI == EPI_BACK) {
                    const float z = g.zprev[o];
                    const float s = sigmoid(z);
                    g.outH[o] = bf16_rne(v * (s * (1.0f + z * (1.0f - s))));
                } else {
                    g.outF[(long long)blockIdx.z * g.M * g.N + o] = v;
                }
            }
        }
    }
}

// Tile shape: the largest of 128 x 128, 128 x 64, 64 x 64 that still gives
// 256 blocks (one per CU); small problems (the 4096-row training batches)
// take the smallest tiles so that more CUs work on them.  Every output element
// accumulates K in the same order whatever the tile, so results are identical.
template <int EPI>
hipError_t gemm(const GemmArgs& g, hipStream_t s, int splits = 1) {
    auto blocks = [&](int bm, int bn) {
        return (long long)((g.N + bn - 1) / bn) * ((g.M + bm - 1) / bm) * splits;
    };
    auto launch = [&](auto kern, int bm, int bn) {
        const dim3 grid((unsigned)((g.N + bn - 1) / bn), (unsigned)((g.M + bm - 1) / bm), (unsigned)splits);
        hipLaunchKernelGGL(kern, grid, dim3(256), 0, s, g);
    };
    if (g.N > 64 && blocks(128, 128) >= 256) launch(gemm_bf16_kernel<EPI, 128, 128>, 128, 128);
    else if (blocks(128, 64) >= 256) launch(gemm_bf16_kernel<EPI, 64, 128>, 128, 64);
    else launch(gemm_bf16_kernel<EPI, 64, 64>, 64, 64);
    return hipGetLastError();
}

// ---- fused inference ------------------------------------------------------
// One block carries 128 rows through every layer of one network: the BF16
// activations stay in LDS (128 x (H + 8) halves, 66 KiB at H = 256, two
// blocks per CU), weights come from L2 in MFMA fragment order, and only the
// FP32 output row goes to HBM.  Each output element accumulates K in steps of 16 in the same
// order as gemm_bf16_kernel, with the same bias, SiLU and rounding, so the
// outputs equal the layer-by-layer GEMM path bit for bit.
constexpr int kFusedRows = 128;
constexpr int kFusedMaxMembers = 8;
constexpr int kFusedMaxLayers = 16;
constexpr int kFusedKin = 16;     // the input layer's K, zero-padded

// Fragment order of a K x N weight matrix (K a multiple of 16): for 32-column
// tile ct and k-step ks, the 64 lanes' B operands of one MFMA lie in 1 KiB,
// lane l holding column 32 ct + l % 32, k = 16 ks + 8 (l / 32) .. + 7.
__host__ __device__ __forceinline__ long long frag_index(int k, int c, int K) {
    const int ct = c >> 5, ks = k >> 4, ln = (c & 31) + 32 * ((k >> 3) & 1);
    return (((long long)ct * (K >> 4) + ks) * 64 + ln) * 8 + (k & 7);
}

// out (Kpad x N, fragment order) = W (K x N row-major BF16), zero for k >= K.
__global__ void frag_kernel(const uint16_t* w, int K, int N, int Kpad, uint16_t* out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)Kpad * N) return;
    const int k = (int)(i / N), c = (int)(i % N);
    out[frag_index(k, c, Kpad)] = k < K ? w[(long long)k * N + c] : (uint16_t)0;
}

struct FusedArgs {
    const float* x;                          // rows x nin, FP32
    int rows, nin, nhh;                      // nhh: hidden-to-hidden layers
    const uint16_t* wf[kFusedMaxMembers];    // fragment-ordered weights per member
    const float* p[kFusedMaxMembers];        // FP32 parameter vector per member
    float* y[kFusedMaxMembers];              // rows x nout outputs per member
    long long boff[kFusedMaxLayers];         // bias offsets in the parameter vector
};

__device__ __forceinline__ float silu(float z) { return z * sigmoid(z); }

// One phase of the pipelined hidden layers (128-row blocks), computed as
// Z^T = W^T X^T: the fragment-ordered weights are the MFMA's A operand, the
// activations in LDS the B operand, so register 4 g + q of a lane holds
// feature w 32 + 8 g + 4 (lane / 32) + q of row lane % 32 of its 32-row tile,
// and the SiLU outputs go back to LDS as 8-byte writes of four features.
// The phase runs the MFMA work of rows [64 hm, 64 hm + 64) of a K -> H layer
// into acc (this wave's 32 features, started from the bias) and, spread over
// its k steps, the epilogue of prev, rows [64 he, 64 he + 64) of the layer
// before (EPI).  Weight fragments are requested 4 k steps ahead, activations
// 1 step ahead.
template <int H, int K, bool EPI>
__device__ __forceinline__ void phase(uint16_t* act, const uint16_t* wl, const float* bias, int hm,
                                      v16f acc[2], int he, const v16f prev[2], int lane, int w) {
    constexpr int LD = H + 8, KS = K / 16, PF = KS < 4 ? KS : 4;
    const v8bf* wv = reinterpret_cast<const v8bf*>(wl);
#pragma unroll
    for (int g = 0; g < 4; ++g)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float b0 = bias[w * 32 + 8 * g + 4 * (lane >> 5) + q];
            acc[0][4 * g + q] = b0;
            acc[1][4 * g + q] = b0;
        }
    auto bload = [&](int ks, int j) {
        return *reinterpret_cast<const v8bf*>(
            &act[((2 * hm + j) * 32 + (lane & 31)) * LD + ks * 16 + 8 * (lane >> 5)]);
    };
    v8bf wq[PF];
#pragma unroll
    for (int i = 0; i < PF; ++i) wq[i] = wv[(w * KS + i) * 64 + lane];
    v8bf b0 = bload(0, 0), b1 = bload(0, 1);
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
        const v8bf av = wq[ks % PF];
        if (ks + PF < KS) wq[ks % PF] = wv[(w * KS + ks + PF) * 64 + lane];
        const v8bf c0 = b0, c1 = b1;
        if (ks + 1 < KS) {
            b0 = bload(ks + 1, 0);
            b1 = bload(ks + 1, 1);
        }
        acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, c0, acc[0], 0, 0, 0);
        acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, c1, acc[1], 0, 0, 0);
        // epilogue share of this k step: (g, j) pairs spread evenly over KS
        if (EPI) {
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                if ((e * KS) / 8 != ks) continue;
                const int g = e >> 1, j = e & 1;
                const int m = w * 32 + 8 * g + 4 * (lane >> 5);
                uint2 v;
                v.x = bf16x2_rne(silu(prev[j][4 * g]), silu(prev[j][4 * g + 1]));
                v.y = bf16x2_rne(silu(prev[j][4 * g + 2]), silu(prev[j][4 * g + 3]));
                *reinterpret_cast<uint2*>(&act[((2 * he + j) * 32 + (lane & 31)) * LD + m]) = v;
            }
        }
    }
}

template <int H>
__device__ __forceinline__ void half_epilogue(uint16_t* act, const v16f acc[2], int hf, int lane, int w) {
    constexpr int LD = H + 8;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const int m = w * 32 + 8 * g + 4 * (lane >> 5);
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            uint2 v;
            v.x = bf16x2_rne(silu(acc[j][4 * g]), silu(acc[j][4 * g + 1]));
            v.y = bf16x2_rne(silu(acc[j][4 * g + 2]), silu(acc[j][4 * g + 3]));
            *reinterpret_cast<uint2*>(&act[((2 * hf + j) * 32 + (lane & 31)) * LD + m]) = v;
        }
    }
}

// Network nin -> H (x 1 + nhh, SiLU) -> NOUT for rows [128 bx, 128 bx + 128)
// of member blockIdx.y, H / 32 waves per block.  In the output layer the
// 32 x 32 tiles (NOUT / 32 feature tiles x 4 row tiles) go round-robin to the
// waves, and each lane writes four outputs of its row per 16-byte store.
template <int H, int NOUT, int NT>
__global__ void __launch_bounds__(2 * H, 2) mlp_infer_kernel(FusedArgs a) {
    constexpr int R = 32 * NT, LD = H + 8, W = H / 32, T = (NOUT / 32) * NT, TPW = (T + W - 1) / W;
    __shared__ __align__(16) uint16_t act[R * LD];
    const int m = blockIdx.y;
    const long long r0 = (long long)blockIdx.x * R;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint16_t* wf = a.wf[m];
    const float* p = a.p[m];
    for (int i = threadIdx.x; i < R * kFusedKin; i += 64 * W) {
        const int r = i / kFusedKin, c = i % kFusedKin;
        const float v = (r0 + r < a.rows && c < a.nin) ? a.x[(r0 + r) * a.nin + c] : 0.0f;
        act[r * LD + c] = bf16_rne(v);
    }
    __syncthreads();
    static_assert(NT == 4, "the hidden-layer pipeline works on two 64-row halves");
    const uint16_t* wl = wf + H * kFusedKin;
    {
        // The hidden layers as a pipeline over the block's two 64-row halves:
        // while a wave's matrix cores take one half of layer l, its vector
        // ALUs apply the SiLU epilogue to the other half (rows that the
        // matrix work does not read); a barrier ends each phase.
        //   A(l): MFMA rows 0-63 of layer l   | epilogue rows 64-127 of l - 1
        //   B(l): MFMA rows 64-127 of layer l | epilogue rows 0-63 of l
        // (layer 0 is the input layer, K = 16)
        v16f lo[2], hi[2];
        phase<H, kFusedKin, false>(act, wf, p + a.boff[0], 0, lo, 1, hi, lane, w);
        __syncthreads();
        phase<H, kFusedKin, false>(act, wf, p + a.boff[0], 1, hi, 0, lo, lane, w);
        half_epilogue<H>(act, lo, 0, lane, w);
        __syncthreads();
        for (int l = 1; l <= a.nhh; ++l) {   // each phase is straight-line code
            const uint16_t* wcur = wf + H * kFusedKin + (long long)(l - 1) * H * H;
            phase<H, H, true>(act, wcur, p + a.boff[l], 0, lo, 1, hi, lane, w);
            __syncthreads();
            phase<H, H, true>(act, wcur, p + a.boff[l], 1, hi, 0, lo, lane, w);
            __syncthreads();
        }
        half_epilogue<H>(act, hi, 1, lane, w);
        __syncthreads();
        wl += (long long)a.nhh * H * H;
    }
    const v8bf* wv = reinterpret_cast<const v8bf*>(wl);
    const float* bias = p + a.boff[a.nhh + 1];
#pragma unroll
    for (int q = 0; q < TPW; ++q) {
        const int t = w + W * q;
        if (t >= T) break;   // uniform per wave
        const int nt = t % NT, mt = t / NT;
        v16f acc;
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[r] = bias[mt * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3)];
#pragma unroll 4
        for (int ks = 0; ks < H / 16; ++ks) {
            const v8bf b = *reinterpret_cast<const v8bf*>(&act[(nt * 32 + (lane & 31)) * LD + ks * 16 + 8 * (lane >> 5)]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wv[(mt * (H / 16) + ks) * 64 + lane], b, acc, 0, 0, 0);
        }
        const long long row = r0 + nt * 32 + (lane & 31);
        if (row < a.rows) {
            float* y = a.y[m] + row * NOUT;
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int c = mt * 32 + 8 * g + 4 * (lane >> 5);
                *reinterpret_cast<float4*>(&y[c]) = float4{acc[4 * g], acc[4 * g + 1], acc[4 * g + 2], acc[4 * g + 3]};
            }
        }
    }
}

// out[i] = sum over s < splits of part[s n + i], in order of s (deterministic).
// Loads go out eight at a time; the additions keep the order of s.
__global__ void split_sum_kernel(const float* part, int splits, long long n, float* out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float acc = 0.0f;
    int s = 0;
    for (; s + 8 <= splits; s += 8) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = part[(long long)(s + u) * n + i];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += v[u];
    }
    for (; s < splits; ++s) acc += part[(long long)s * n + i];
    out[i] = acc;
}

// ---- elementwise kernels ---------------------------------------------------
__global__ void to_bf16_kernel(const float* in, uint16_t* out, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = bf16_rne(in[i]);
}

// Rows idx[r] of the data set (nin and nout columns) into the batch buffers.
__global__ void gather_kernel(const float* xd, const float* td, const uint32_t* idx, int nrow,
                              int nin, int nout, uint16_t* xb, float* tb) {
    const int r = blockIdx.x;
    if (r >= nrow) return;
    const long long s = idx[r];
    for (int c = threadIdx.x; c < nin; c += blockDim.x) xb[(long long)r * nin + c] = bf16_rne(xd[s * nin + c]);
    for (int c = threadIdx.x; c < nout; c += blockDim.x) tb[(long long)r * nout + c] = td[s * nout + c];
}

// Mini-batch rows drawn with replacement: Philox(step, r) mod ndata.
__global__ void batch_index_kernel(uint32_t* idx, int nrow, uint32_t ndata, uint32_t step,
                                   uint32_t k0, uint32_t k1) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrow) return;
    const U4 u = philox4x32_10((uint32_t)r, step, 0x6d6c70u, 0u, k0, k1);
    idx[r] = (uint32_t)(((unsigned long long)u.a * ndata) >> 32);
}

// Mean squared error over n x nout: per-block partial sums of e^2 and the
// output gradient 2 e / (n nout) rounded to BF16.
__global__ void __launch_bounds__(256) mse_kernel(const float* y, const float* t, long long count,
                                                  float scale, uint16_t* g, float* part) {
    __shared__ float red[256];
    float acc = 0.0f;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < count; i += (long long)gridDim.x * 256) {
        const float e = y[i] - t[i];
        acc += e * e;
        g[i] = bf16_rne(2.0f * e * scale);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Ensemble output transform, one thread per row: each member's spectrum
// s_i = max(10^y_i - floor, 0), their mean, and the mean L1 distance of the
// members from it.
struct EnsembleArgs {
    const float* y[8];
    int members, nout;
    float floor;
    float* spec;
    float* spread;
};

// 10^y from the hardware exp2 (v_exp_f32, about 1 ulp): for the outputs'
// range (|y| <= 6) the relative error stays below 1e-6.
__device__ __forceinline__ float pow10f(float y) { return __builtin_amdgcn_exp2f(y * 3.32192809488736235f); }

// Sixteen lanes per row, four consecutive columns per lane and pass (16-byte
// loads and stores when nout % 4 == 0); a lane's member distances add up in
// column order, then the 16 lanes' sums by a fixed butterfly.
__device__ __forceinline__ void ensemble_column(const EnsembleArgs& e, long long o, float inv, float& dist) {
    float sv[8];
    float mean = 0.0f;
#pragma unroll
    for (int i = 0; i < 8; ++i)
        if (i < e.members) {
            sv[i] = fmaxf(pow10f(e.y[i][o]) - e.floor, 0.0f);
            mean += sv[i];
        }
    mean *= inv;
#pragma unroll
    for (int i = 0; i < 8; ++i)
        if (i < e.members) dist += fabsf(sv[i] - mean);
    e.spec[o] = mean;
}

__global__ void __launch_bounds__(256) ensemble_kernel(EnsembleArgs e, int rows) {
    const int l16 = threadIdx.x & 15;
    const long long r = (long long)blockIdx.x * 16 + (threadIdx.x >> 4);
    if (r >= rows) return;   // uniform per 16-lane group, the shuffle width below
    const float inv = 1.0f / (float)e.members;
    const long long base = r * e.nout;
    float dist = 0.0f;
    for (int c = 4 * l16; c < e.nout; c += 64) {
        if ((e.nout & 3) == 0) {
            float4 yv[8];
#pragma unroll
            for (int i = 0; i < 8; ++i)
                if (i < e.members) yv[i] = *reinterpret_cast<const float4*>(e.y[i] + base + c);
            float4 mv;
            float* mp = &mv.x;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                float sv[8];
                float mean = 0.0f;
#pragma unroll
                for (int i = 0; i < 8; ++i)
                    if (i < e.members) {
                        sv[i] = fmaxf(pow10f((&yv[i].x)[k]) - e.floor, 0.0f);
                        mean += sv[i];
                    }
                mean *= inv;
#pragma unroll
                for (int i = 0; i < 8; ++i)
                    if (i < e.members) dist += fabsf(sv[i] - mean);
                mp[k] = mean;
            }
            *reinterpret_cast<float4*>(e.spec + base + c) = mv;
        } else {
            for (int k = c; k < min(c + 4, e.nout); ++k) ensemble_column(e, base + k, inv, dist);
        }
    }
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) dist += __shfl_xor(dist, o, 16);
    if (l16 == 0) e.spread[r] = dist * inv;
}

// out[0] = scale * sum(part[0 .. np)), one block.
__global__ void __launch_bounds__(256) sum_scale_kernel(const float* part, int np, float scale,
                                                        float* out) {
    __shared__ float red[256];
    float acc = 0.0f;
    for (int i = threadIdx.x; i < np; i += 256) acc += part[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0] * scale;
}

__global__ void zero_kernel(float* a, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.0f;
}

// Column sums of a BF16 M x N matrix (bias gradients), one block per column.
// Partial column sums of a rows x N BF16 matrix: block (bx, by) sums columns
// [64 bx, 64 bx + 64) over rows [by chunk, (by + 1) chunk).  A thread takes
// eight columns (one 16-byte load per row) and every 32nd row; the 32 row
// phases are then added in a fixed order.  part[by N + c].
constexpr int kColParts = 16;
__global__ void __launch_bounds__(256) colsum_part_kernel(const uint16_t* gmat, int M, int N,
                                                          int chunk, float* part) {
    __shared__ float red[32][65];
    const int cl = (threadIdx.x & 7) * 8, c0 = blockIdx.x * 64 + cl, ph = threadIdx.x >> 3;
    const int r0 = blockIdx.y * chunk, r1 = min(M, r0 + chunk);
    const bool vec = (N % 8) == 0 && c0 + 8 <= N;
    float acc[8] = {};
#pragma unroll 8
    for (int r = r0 + ph; r < r1; r += 32) {
        if (vec) {
            const uint4 v = *reinterpret_cast<const uint4*>(gmat + (long long)r * N + c0);
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                acc[2 * u] += __uint_as_float(w[u] << 16);
                acc[2 * u + 1] += __uint_as_float(w[u] & 0xFFFF0000u);
            }
        } else {
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (c0 + u < N) acc[u] += bf16_to_f(gmat[(long long)r * N + c0 + u]);
        }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) red[ph][cl + u] = acc[u];
    __syncthreads();
    if (threadIdx.x < 64) {
        const int c = blockIdx.x * 64 + threadIdx.x;
        float t = 0.0f;
        for (int q = 0; q < 32; ++q) t += red[q][threadIdx.x];
        if (c < N) part[(long long)blockIdx.y * N + c] = t;
    }
}

// State of a device-side training loop: t counts the AdamW steps taken; the
// loop's kernels read it, so one captured step replays unchanged.
struct TrainState {
    uint32_t t, t0, nsteps, pad;
    float lr0, lr_min, b1, b2;
};

__global__ void batch_index_dev_kernel(uint32_t* idx, int nrow, uint32_t ndata,
                                       const TrainState* ts, uint32_t k0, uint32_t k1) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrow) return;
    const U4 u = philox4x32_10((uint32_t)r, ts->t - ts->t0, 0x6d6c70u, 0u, k0, k1);
    idx[r] = (uint32_t)(((unsigned long long)u.a * ndata) >> 32);
}

// loss_hist[t - t0] = scale * sum(part[0 .. np)), one block.
__global__ void __launch_bounds__(256) loss_dev_kernel(const float* part, int np, float scale,
                                                       float* loss_hist, const TrainState* ts) {
    __shared__ float red[256];
    float acc = 0.0f;
    for (int i = threadIdx.x; i < np; i += 256) acc += part[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_hist[ts->t - ts->t0] = red[0] * scale;
}

// AdamW step t + 1 with the learning rate of the loop's cosine schedule.
__global__ void adamw_dev_kernel(float* p, const float* gr, float* m, float* v, uint16_t* pb,
                                 long long n, const TrainState* ts, float eps, float wd) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = ts->t - ts->t0;
    const float f = ts->nsteps > 1 ? (float)s / (float)(ts->nsteps - 1) : 0.0f;
    const float lr = ts->lr_min + 0.5f * (ts->lr0 - ts->lr_min) * (1.0f + cosf(3.14159265358979f * f));
    const float tt = (float)(ts->t + 1);
    const float c1 = 1.0f - powf(ts->b1, tt), c2 = 1.0f - powf(ts->b2, tt);
    const float g = gr[i];
    const float mi = ts->b1 * m[i] + (1.0f - ts->b1) * g;
    const float vi = ts->b2 * v[i] + (1.0f - ts->b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float upd = (mi / c1) / (sqrtf(vi / c2) + eps);
    const float pi = p[i] - lr * (upd + wd * p[i]);
    p[i] = pi;
    pb[i] = bf16_rne(pi);
}

__global__ void train_tick_kernel(TrainState* ts) {
    if (threadIdx.x == 0 && blockIdx.x == 0) ts->t += 1u;
}

// AdamW (decoupled weight decay) with bias correction; refreshes the BF16 copy.
__global__ void adamw_kernel(float* p, const float* gr, float* m, float* v, uint16_t* pb,
                             long long n, float lr, float b1, float b2, float eps, float wd,
                             float c1, float c2) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = gr[i];
    const float mi = b1 * m[i] + (1.0f - b1) * g;
    const float vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float upd = (mi / c1) / (sqrtf(vi / c2) + eps);
    const float pi = p[i] - lr * (upd + wd * p[i]);
    p[i] = pi;
    pb[i] = bf16_rne(pi);
}

}  // namespace mlp
}  // namespace sp

// ---------------------------------------------------------------------------
// Host side: the network context and the extern "C" entry points.
namespace {

struct Mlp {
    sp_mlp_config cfg{};
    hipStream_t st = nullptr;
    int L = 0;                          // linear layers
    std::vector<int> K, N;              // inputs / outputs of layer l
    std::vector<long long> woff, boff;  // offsets of W_l, b_l in the parameter vector
    long long P = 0;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    uint16_t* pb = nullptr;
    int bmax = 0, nmax = 0;
    std::vector<uint16_t*> xh;          // X_l (bf16), l = 0 .. L-1
    std::vector<float*> z;              // Z_l (fp32), l = 0 .. L-2
    float* y = nullptr;                 // network output (fp32)
    float* t = nullptr;                 // targets of the batch (fp32)
    uint16_t* gh[2] = {nullptr, nullptr};
    float* part = nullptr;
    float* loss_hist = nullptr;
    long long loss_cap = 0;
    float* xin = nullptr;               // host-supplied input rows (fp32)
    float* xd = nullptr;                // training data set on the device
    float* td = nullptr;
    long long nd = 0;
    uint32_t* bidx = nullptr;
    float* wpart = nullptr;             // split-K partial weight gradients
    hipEvent_t ev[2] = {nullptr, nullptr};
    double fwd_ms = 0.0;                // device time of the last forward's layer kernels
    sp::mlp::TrainState* ts = nullptr;  // device-side training loop state
    float* ens_spec = nullptr;          // ensemble spectra and spreads (batch_max rows)
    float* ens_spread = nullptr;
    uint16_t* wf = nullptr;             // fragment-ordered BF16 weights (fused inference)
    uint64_t pver = 1, fver = 0;        // versions of the BF16 weights and of wf
    float* ey = nullptr;                // outputs of device-array ensemble chunks
    long long ey_rows = 0;
    std::vector<hipEvent_t> tev;        // per-chunk timing events of the fused launches
    long long step = 0;                 // AdamW step count
};

constexpr int kPartBlocks = 256;
constexpr int kMaxSplits = 32;

void dfree(void* q) {
    if (q) (void)hipFree(q);
}

void destroy(Mlp* n) {
    if (!n) return;
    if (n->st) (void)hipStreamSynchronize(n->st);
    dfree(n->p); dfree(n->g); dfree(n->m); dfree(n->v); dfree(n->pb);
    for (auto* q : n->xh) dfree(q);
    for (auto* q : n->z) dfree(q);
    dfree(n->y); dfree(n->t); dfree(n->gh[0]); dfree(n->gh[1]); dfree(n->part);
    dfree(n->loss_hist); dfree(n->xin); dfree(n->xd); dfree(n->td); dfree(n->bidx);
    dfree(n->wpart); dfree(n->ens_spec); dfree(n->ens_spread); dfree(n->ts); dfree(n->wf);
    dfree(n->ey);
    for (auto& e : n->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : n->tev)
        if (e) (void)hipEventDestroy(e);
    if (n->st) (void)hipStreamDestroy(n->st);
    delete n;
}

#define MLP_HIP(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return sp::set_error(SP_ERR_HIP, std::string(#expr) + " failed: " + hipGetErrorString(e_)); \
    } while (0)

// Deterministic initial parameters: U(-1/sqrt(K), 1/sqrt(K)) for W_l and b_l
// from a splitmix64 stream of the configured seed.
void init_params(const Mlp* n, std::vector<float>& h) {
    uint64_t s = n->cfg.seed;
    auto next = [&s]() {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    h.assign((size_t)n->P, 0.0f);
    for (int l = 0; l < n->L; ++l) {
        const double a = 1.0 / std::sqrt((double)n->K[l]);
        const long long nw = (long long)n->K[l] * n->N[l];
        for (long long i = 0; i < nw; ++i) h[(size_t)(n->woff[l] + i)] = (float)((2.0 * next() - 1.0) * a);
        for (int i = 0; i < n->N[l]; ++i) h[(size_t)(n->boff[l] + i)] = (float)((2.0 * next() - 1.0) * a);
    }
}

unsigned blocks(long long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// Forward pass of rows [0, rows) of X_0 (already in xh[0]).
int forward(Mlp* n, int rows, bool train = true) {
    for (int l = 0; l < n->L; ++l) {
        sp::mlp::GemmArgs g{};
        g.A = n->xh[l];
        g.lda = n->K[l];
        g.B = n->pb + n->woff[l];
        g.ldb = n->N[l];
        g.M = rows; g.N = n->N[l]; g.K = n->K[l];
        g.bias = n->p + n->boff[l];
        if (l + 1 < n->L) {
            g.outF = n->z[l];
            g.outH = n->xh[l + 1];
            if (train) MLP_HIP(sp::mlp::gemm<sp::mlp::EPI_HIDDEN>(g, n->st));
            else MLP_HIP(sp::mlp::gemm<sp::mlp::EPI_INFER>(g, n->st));
        } else {
            g.outF = n->y;
            MLP_HIP(sp::mlp::gemm<sp::mlp::EPI_LINEAR>(g, n->st));
        }
    }
    return SP_OK;
}

// Loss, gradients of all parameters into n->g; the mean squared error goes to
// loss_out (device), or with a training state to loss_out[t - t0].
int backward(Mlp* n, int rows, float* loss_out, const sp::mlp::TrainState* ts = nullptr) {
    const long long cnt = (long long)rows * n->cfg.nout;
    const float scale = 1.0f / (float)cnt;
    hipLaunchKernelGGL(sp::mlp::mse_kernel, dim3(kPartBlocks), dim3(256), 0, n->st, n->y, n->t, cnt,
                       scale, n->gh[0], n->part);
    MLP_HIP(hipGetLastError());
    if (ts)
        hipLaunchKernelGGL(sp::mlp::loss_dev_kernel, dim3(1), dim3(256), 0, n->st, n->part,
                           kPartBlocks, scale, loss_out, ts);
    else
        hipLaunchKernelGGL(sp::mlp::sum_scale_kernel, dim3(1), dim3(256), 0, n->st, n->part,
                           kPartBlocks, scale, loss_out);
    MLP_HIP(hipGetLastError());
    int cur = 0;
    for (int l = n->L - 1; l >= 0; --l) {
        // weight gradient: K = rows is long and M x N small, so K is split
        // over blocks into partial products summed in a fixed order
        sp::mlp::GemmArgs w{};
        w.A = n->xh[l]; w.lda = n->K[l]; w.transA = 1;
        w.B = n->gh[cur]; w.ldb = n->N[l];
        w.M = n->K[l]; w.N = n->N[l]; w.K = rows;
        const int bn = w.N > 64 ? 128 : 64;
        const int tiles = ((w.M + 127) / 128) * ((w.N + bn - 1) / bn);
        int splits = std::max(1, std::min(kMaxSplits, 256 / std::max(1, tiles)));
        splits = std::min(splits, (rows + 127) / 128);
        w.kchunk = ((rows + splits - 1) / splits + 31) / 32 * 32;
        splits = (rows + w.kchunk - 1) / w.kchunk;
        const long long mn = (long long)w.M * w.N;
        w.outF = n->wpart;
        MLP_HIP(sp::mlp::gemm<sp::mlp::EPI_PLAIN>(w, n->st, splits));
        hipLaunchKernelGGL(sp::mlp::split_sum_kernel, dim3(blocks(mn)), dim3(256), 0, n->st,
                           n->wpart, splits, mn, n->g + n->woff[l]);
        MLP_HIP(hipGetLastError());
        // bias gradient: column sums of the output gradient in kColParts row
        // chunks, then summed in order of the chunk
        const int chunk = (rows + sp::mlp::kColParts - 1) / sp::mlp::kColParts;
        float* bpart = n->wpart + (long long)splits * mn;
        hipLaunchKernelGGL(sp::mlp::colsum_part_kernel,
                           dim3((unsigned)((n->N[l] + 63) / 64), (unsigned)sp::mlp::kColParts), dim3(256),
                           0, n->st, n->gh[cur], rows, n->N[l], chunk, bpart);
        MLP_HIP(hipGetLastError());
        hipLaunchKernelGGL(sp::mlp::split_sum_kernel, dim3(blocks(n->N[l])), dim3(256), 0, n->st,
                           bpart, sp::mlp::kColParts, (long long)n->N[l], n->g + n->boff[l]);
        MLP_HIP(hipGetLastError());
        if (l > 0) {
            sp::mlp::GemmArgs d{};
            d.A = n->gh[cur]; d.lda = n->N[l];
            d.B = n->pb + n->woff[l]; d.ldb = n->N[l]; d.transB = 1;
            d.M = rows; d.N = n->K[l]; d.K = n->N[l];
            d.zprev = n->z[l - 1];
            d.outH = n->gh[1 - cur];
            MLP_HIP(sp::mlp::gemm<sp::mlp::EPI_BACK>(d, n->st));
            cur = 1 - cur;
        }
    }
    return SP_OK;
}

int adamw(Mlp* n, double lr) {
    n->step += 1;
    n->pver += 1;
    const double c1 = 1.0 - std::pow(n->cfg.beta1, (double)n->step);
    const double c2 = 1.0 - std::pow(n->cfg.beta2, (double)n->step);
    hipLaunchKernelGGL(sp::mlp::adamw_kernel, dim3(blocks(n->P)), dim3(256), 0, n->st, n->p, n->g,
                       n->m, n->v, n->pb, n->P, (float)lr, (float)n->cfg.beta1,
                       (float)n->cfg.beta2, (float)n->cfg.eps, (float)n->cfg.weight_decay,
                       (float)c1, (float)c2);
    MLP_HIP(hipGetLastError());
    return SP_OK;
}

// Host rows -> xh[0] (bf16) and, when given, targets -> n->t.
// x and t may be host or device pointers.
int load_rows(Mlp* n, const float* x, const float* t, int rows) {
    const long long nx = (long long)rows * n->cfg.nin;
    MLP_HIP(hipMemcpyAsync(n->xin, x, sizeof(float) * (size_t)nx, hipMemcpyDefault, n->st));
    hipLaunchKernelGGL(sp::mlp::to_bf16_kernel, dim3(blocks(nx)), dim3(256), 0, n->st, n->xin,
                       n->xh[0], nx);
    MLP_HIP(hipGetLastError());
    if (t)
        MLP_HIP(hipMemcpyAsync(n->t, t, sizeof(float) * (size_t)rows * n->cfg.nout,
                               hipMemcpyDefault, n->st));
    return SP_OK;
}

Mlp* live(void* q, int& rc) {
    if (!q) {
        rc = sp::set_error(SP_ERR_ARG, "network is NULL");
        return nullptr;
    }
    rc = SP_OK;
    return static_cast<Mlp*>(q);
}

// Shapes the fused inference kernel is built for; SP_MLP_UNFUSED=1 sends
// inference through the layer-by-layer GEMMs instead.
bool fused_ok(const Mlp* n) {
    const char* env = std::getenv("SP_MLP_UNFUSED");
    if (env && env[0] == '1') return false;
    const sp_mlp_config& c = n->cfg;
    return c.nin <= sp::mlp::kFusedKin && (c.nhidden == 128 || c.nhidden == 256) &&
           (c.nout == 32 || c.nout == 64 || c.nout == 128) && c.nlayers >= 1 &&
           n->L <= sp::mlp::kFusedMaxLayers;
}

bool same_shape(const Mlp* a, const Mlp* b) {
    return a->cfg.nin == b->cfg.nin && a->cfg.nhidden == b->cfg.nhidden &&
           a->cfg.nlayers == b->cfg.nlayers && a->cfg.nout == b->cfg.nout;
}

// Rewrites n->wf (fragment order, input layer padded to K = 16) from the BF16
// weights on stream s, unless they have not changed since the last rewrite
// (pver counts the updates of the BF16 weights).
int refresh_frag(Mlp* n, hipStream_t s) {
    if (n->wf && n->fver == n->pver) return SP_OK;
    long long total = 0;
    for (int l = 0; l < n->L; ++l) total += (long long)(l == 0 ? sp::mlp::kFusedKin : n->K[l]) * n->N[l];
    if (!n->wf) MLP_HIP(hipMalloc(&n->wf, 2 * (size_t)total));
    long long off = 0;
    for (int l = 0; l < n->L; ++l) {
        const int kpad = l == 0 ? sp::mlp::kFusedKin : n->K[l];
        const long long cnt = (long long)kpad * n->N[l];
        hipLaunchKernelGGL(sp::mlp::frag_kernel, dim3(blocks(cnt)), dim3(256), 0, s,
                           n->pb + n->woff[l], n->K[l], n->N[l], kpad, n->wf + off);
        MLP_HIP(hipGetLastError());
        off += cnt;
    }
    n->fver = n->pver;
    return SP_OK;
}

template <int H, int NOUT>
void launch_fused(const sp::mlp::FusedArgs& a, int members, hipStream_t s) {
    // 128-row tiles (64-row tiles, more blocks per CU, measured 12 % slower)
    const dim3 grid((unsigned)((a.rows + sp::mlp::kFusedRows - 1) / sp::mlp::kFusedRows), (unsigned)members);
    hipLaunchKernelGGL((sp::mlp::mlp_infer_kernel<H, NOUT, sp::mlp::kFusedRows / 32>), grid, dim3(2 * H), 0, s, a);
}

// y[i] (rows x nout; ms[i]->y when y is NULL) = network i (x), all members
// in one launch on s.  Needs fused_ok and equal shapes, and every member's wf
// refreshed.
int fused_forward(Mlp* const* ms, int members, const float* x, int rows, hipStream_t s,
                  float* const* y = nullptr) {
    const Mlp* h = ms[0];
    sp::mlp::FusedArgs a{};
    a.x = x;
    a.rows = rows;
    a.nin = h->cfg.nin;
    a.nhh = h->L - 2;
    for (int i = 0; i < members; ++i) {
        a.wf[i] = ms[i]->wf;
        a.p[i] = ms[i]->p;
        a.y[i] = y ? y[i] : ms[i]->y;
    }
    for (int l = 0; l < h->L; ++l) a.boff[l] = h->boff[l];
    if (rows == 0) return SP_OK;
    const int H = h->cfg.nhidden, NO = h->cfg.nout;
    if (H == 256 && NO == 64) launch_fused<256, 64>(a, members, s);
    else if (H == 256 && NO == 32) launch_fused<256, 32>(a, members, s);
    else if (H == 256 && NO == 128) launch_fused<256, 128>(a, members, s);
    else if (H == 128 && NO == 64) launch_fused<128, 64>(a, members, s);
    else if (H == 128 && NO == 32) launch_fused<128, 32>(a, members, s);
    else launch_fused<128, 128>(a, members, s);
    MLP_HIP(hipGetLastError());
    return SP_OK;
}

int ensure_events(Mlp* n, size_t count) {
    while (n->tev.size() < count) {
        hipEvent_t e = nullptr;
        MLP_HIP(hipEventCreate(&e));
        n->tev.push_back(e);
    }
    return SP_OK;
}

// Ensemble spectra of rows of x (host or device pointers throughout; dev
// selects which).  Chunks of batch_max rows; with the fused kernel all
// members of a chunk run in one launch on the first member's stream.
// True for memory that kernels may read: device or managed allocations.
bool device_memory(const void* q) {
    hipPointerAttribute_t at{};
    if (hipPointerGetAttributes(&at, q) != hipSuccess) {
        (void)hipGetLastError();   // an unknown host pointer; clear the error
        return false;
    }
    return at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged;
}

int ensemble(void* const* nets, int32_t members, const float* x, int64_t rows, double floor,
             float* spec, float* spread, bool dev) {
    if (!nets || members < 1 || members > 8) return sp::set_error(SP_ERR_ARG, "members must be in [1, 8]");
    if (rows < 0 || (rows > 0 && (!x || !spec || !spread))) return sp::set_error(SP_ERR_ARG, "bad rows or NULL buffers");
    if (dev && rows > 0 && !(device_memory(x) && device_memory(spec) && device_memory(spread)))
        return sp::set_error(SP_ERR_ARG, "x, spec and spread must be device memory");
    std::vector<Mlp*> ms;
    int rc;
    for (int i = 0; i < members; ++i) {
        Mlp* n = live(nets[i], rc);
        if (!n) return rc;
        if (n->cfg.nin != static_cast<Mlp*>(nets[0])->cfg.nin || n->cfg.nout != static_cast<Mlp*>(nets[0])->cfg.nout)
            return sp::set_error(SP_ERR_ARG, "ensemble members differ in nin or nout");
        ms.push_back(n);
    }
    Mlp* h = ms[0];
    int chunk = h->bmax;
    bool fused = true;
    for (auto* n : ms) {
        chunk = std::min(chunk, n->bmax);
        fused = fused && fused_ok(n) && same_shape(n, h);
    }
    const int nin = h->cfg.nin, nout = h->cfg.nout;
    if (!dev && !h->ens_spec) {
        MLP_HIP(hipMalloc(&h->ens_spec, 4 * (size_t)h->bmax * nout));
        MLP_HIP(hipMalloc(&h->ens_spread, 4 * (size_t)h->bmax));
    }
    // device arrays in and out: chunks of up to 2^20 rows into the members'
    // own output buffers (fewer, fuller launches than batch_max-row chunks)
    std::vector<float*> ys((size_t)members, nullptr);
    for (int i = 0; i < members; ++i) ys[(size_t)i] = ms[(size_t)i]->y;
    if (dev && fused && rows > 0) {
        chunk = (int)std::min<int64_t>(rows, (int64_t)1 << 20);
        for (int i = 0; i < members; ++i) {
            Mlp* n = ms[(size_t)i];
            if (n->ey_rows < chunk) {
                dfree(n->ey);
                n->ey = nullptr;
                n->ey_rows = 0;
                MLP_HIP(hipMalloc(&n->ey, 4 * (size_t)chunk * nout));
                n->ey_rows = chunk;
            }
            ys[(size_t)i] = n->ey;
        }
    }
    const int64_t nchunks = (rows + chunk - 1) / chunk;
    if (fused) {
        for (auto* n : ms) {
            MLP_HIP(hipStreamSynchronize(n->st));
            rc = refresh_frag(n, h->st);
            if (rc != SP_OK) return rc;
        }
        rc = ensure_events(h, 2 * (size_t)nchunks);
        if (rc != SP_OK) return rc;
    }
    for (int64_t c = 0; c < nchunks; ++c) {
        const int64_t r0 = c * chunk;
        const int r = (int)std::min<int64_t>(chunk, rows - r0);
        if (fused) {
            const float* xc = x + r0 * nin;
            if (!dev) {
                MLP_HIP(hipMemcpyAsync(h->xin, xc, 4 * (size_t)r * nin, hipMemcpyHostToDevice, h->st));
                xc = h->xin;
            }
            MLP_HIP(hipEventRecord(h->tev[(size_t)(2 * c)], h->st));
            rc = fused_forward(ms.data(), members, xc, r, h->st, ys.data());
            if (rc != SP_OK) return rc;
            MLP_HIP(hipEventRecord(h->tev[(size_t)(2 * c + 1)], h->st));
        } else {
            for (auto* n : ms) {
                rc = load_rows(n, x + r0 * nin, nullptr, r);
                if (rc == SP_OK) rc = forward(n, r, false);
                if (rc != SP_OK) return rc;
                MLP_HIP(hipStreamSynchronize(n->st));
            }
        }
        sp::mlp::EnsembleArgs e{};
        for (int i = 0; i < members; ++i) e.y[i] = ys[(size_t)i];
        e.members = members;
        e.nout = nout;
        e.floor = (float)floor;
        e.spec = dev ? spec + r0 * nout : h->ens_spec;
        e.spread = dev ? spread + r0 : h->ens_spread;
        hipLaunchKernelGGL(sp::mlp::ensemble_kernel, dim3(blocks(r, 16)), dim3(256), 0, h->st, e, r);
        MLP_HIP(hipGetLastError());
        if (!dev) {
            MLP_HIP(hipMemcpyAsync(spec + r0 * nout, h->ens_spec, 4 * (size_t)r * nout, hipMemcpyDeviceToHost, h->st));
            MLP_HIP(hipMemcpyAsync(spread + r0, h->ens_spread, 4 * (size_t)r, hipMemcpyDeviceToHost, h->st));
            MLP_HIP(hipStreamSynchronize(h->st));
        }
    }
    MLP_HIP(hipStreamSynchronize(h->st));
    h->fwd_ms = 0.0;
    if (fused)
        for (int64_t c = 0; c < nchunks; ++c) {
            float ms_ = 0.f;
            MLP_HIP(hipEventElapsedTime(&ms_, h->tev[(size_t)(2 * c)], h->tev[(size_t)(2 * c + 1)]));
            h->fwd_ms += ms_;
        }
    return SP_OK;
}

}  // namespace

extern "C" {

int sp_mlp_create(const sp_mlp_config* cfg, void** out) {
    if (!cfg || !out) return sp::set_error(SP_ERR_ARG, "cfg or out is NULL");
    *out = nullptr;
    if (cfg->nin < 1 || cfg->nin > 4096 || cfg->nout < 1 || cfg->nout > 4096 || cfg->nhidden < 1 ||
        cfg->nhidden > 4096 || cfg->nlayers < 0 || cfg->nlayers > 64)
        return sp::set_error(SP_ERR_ARG, "layer sizes out of range");
    if (cfg->batch_max < 1 || cfg->batch_max > (1 << 21))
        return sp::set_error(SP_ERR_ARG, "batch_max must be in [1, 2^21]");
    if (cfg->reserved0 != 0) return sp::set_error(SP_ERR_ARG, "reserved0 must be 0");
    if (!(cfg->beta1 >= 0 && cfg->beta1 < 1 && cfg->beta2 >= 0 && cfg->beta2 < 1 && cfg->eps > 0 &&
          cfg->weight_decay >= 0))
        return sp::set_error(SP_ERR_ARG, "AdamW parameters out of range");
    Mlp* n = new Mlp();
    n->cfg = *cfg;
    n->L = cfg->nlayers + 1;
    for (int l = 0; l < n->L; ++l) {
        n->K.push_back(l == 0 ? cfg->nin : cfg->nhidden);
        n->N.push_back(l + 1 == n->L ? cfg->nout : cfg->nhidden);
    }
    for (int l = 0; l < n->L; ++l) {
        n->woff.push_back(n->P);
        n->P += (long long)n->K[l] * n->N[l];
        n->boff.push_back(n->P);
        n->P += n->N[l];
    }
    n->bmax = cfg->batch_max;
    int wmax = std::max(cfg->nin, std::max(cfg->nhidden, cfg->nout));
    int rc = SP_OK;
    auto fail_all = [&](hipError_t e, const char* what) {
        rc = sp::set_error(SP_ERR_NOMEM, std::string(what) + ": " + hipGetErrorString(e));
        destroy(n);
        return rc;
    };
    hipError_t e = hipStreamCreate(&n->st);
    if (e != hipSuccess) return fail_all(e, "hipStreamCreate");
    for (auto& ev : n->ev)
        if (e == hipSuccess) e = hipEventCreate(&ev);
    if (e != hipSuccess) return fail_all(e, "hipEventCreate");
    const size_t P = (size_t)n->P, B = (size_t)n->bmax;
    e = hipMalloc(&n->p, 4 * P);
    if (e == hipSuccess) e = hipMalloc(&n->g, 4 * P);
    if (e == hipSuccess) e = hipMalloc(&n->m, 4 * P);
    if (e == hipSuccess) e = hipMalloc(&n->v, 4 * P);
    if (e == hipSuccess) e = hipMalloc(&n->pb, 2 * P);
    n->xh.assign((size_t)n->L, nullptr);
    n->z.assign((size_t)std::max(0, n->L - 1), nullptr);
    for (int l = 0; l < n->L && e == hipSuccess; ++l) e = hipMalloc(&n->xh[(size_t)l], 2 * B * (size_t)n->K[l]);
    for (int l = 0; l + 1 < n->L && e == hipSuccess; ++l) e = hipMalloc(&n->z[(size_t)l], 4 * B * (size_t)n->N[l]);
    if (e == hipSuccess) e = hipMalloc(&n->y, 4 * B * (size_t)cfg->nout);
    if (e == hipSuccess) e = hipMalloc(&n->t, 4 * B * (size_t)cfg->nout);
    if (e == hipSuccess) e = hipMalloc(&n->gh[0], 2 * B * (size_t)wmax);
    if (e == hipSuccess) e = hipMalloc(&n->gh[1], 2 * B * (size_t)wmax);
    if (e == hipSuccess) e = hipMalloc(&n->part, 4 * (size_t)kPartBlocks);
    if (e == hipSuccess) e = hipMalloc(&n->xin, 4 * B * (size_t)cfg->nin);
    if (e == hipSuccess) e = hipMalloc(&n->bidx, 4 * B);
    long long wmaxmn = 0;
    for (int l = 0; l < n->L; ++l) wmaxmn = std::max(wmaxmn, (long long)n->K[l] * n->N[l]);
    if (e == hipSuccess)
        e = hipMalloc(&n->wpart, 4 * ((size_t)kMaxSplits * (size_t)wmaxmn + (size_t)sp::mlp::kColParts * wmax));
    if (e == hipSuccess) e = hipMalloc(&n->loss_hist, 4 * 16);
    if (e != hipSuccess) return fail_all(e, "device allocation for the network");
    n->loss_cap = 16;
    std::vector<float> h;
    init_params(n, h);
    rc = sp_mlp_set_params(n, h.data());
    if (rc != SP_OK) {
        destroy(n);
        return rc;
    }
    *out = n;
    return SP_OK;
}

int sp_mlp_destroy(void* q) {
    destroy(static_cast<Mlp*>(q));
    return SP_OK;
}

int sp_mlp_num_params(void* q, int64_t* np) {
    int rc;
    Mlp* n = live(q, rc);
    if (!n) return rc;
    if (!np) return sp::set_error(SP_ERR_ARG, "np is NULL");
    *np = n->P;
    return SP_OK;
}

int sp_mlp_get_params(void* q, float* out) {
    int rc;
    Mlp* n = live(q, rc);
    if (!n) return rc;
    if (!out) return sp::set_error(SP_ERR_ARG, "out is NULL");
    MLP_HIP(hipMemcpyAsync(out, n->p, 4 * (size_t)n->P, hipMemcpyDeviceToHost, n->st));
    MLP_HIP(hipStreamSynchronize(n->st));
    return SP_OK;
}

int sp_mlp_set_params(void* q, const float* in) {
    int rc;
    Mlp* n = live(q, rc);
    if (!n) return rc;
    if (!in) return sp::set_error(SP_ERR_ARG, "in is NULL");
    n->pver += 1;
    MLP_HIP(hipMemcpyAsync(n->p, in, 4 * (size_t)n->P, hipMemcpyHostToDevice, n->st));
    hipLaunchKernelGGL(sp::mlp::to_bf16_kernel, dim3(blocks(n->P)), dim3(256), 0, n->st, n->p, n->pb,
                       n->P);
    MLP_HIP(hipGetLastError());
    hipLaunchKernelGGL(sp::mlp::zero_kernel, dim3(blocks(n->P)), dim3(256), 0, n->st, n->m, n->P);
    hipLaunchKernelGGL(sp::mlp::zero_kernel, dim3(blocks(n->P)), dim3(256), 0, n->st, n->v, n->P);
    MLP_HIP(hipGetLastError());
    MLP_HIP(hipStreamSynchronize(n->st));
    n->step = 0;
    return SP_OK;
}

int sp_mlp_forward(void* q, const float* x, int64_t rows, float* y) {
    int rc;
    Mlp* n = live(q, rc);
    if (!n) return rc;
    if (rows < 0 || (rows > 0 && (!x || !y))) return sp::set_error(SP_ERR_ARG, "bad rows or NULL buffers");
    n->fwd_ms = 0.0;
    const bool fused = fused_ok(n);
    if (fused) {
        rc = refresh_frag(n, n->st);
        if (rc != SP_OK) return rc;
    }
    for (int64_t r0 = 0; r0 < rows; r0 += n->bmax) {
        const int r = (int)std::min<int64_t>(n->bmax, rows - r0);
        rc = load_rows(n, x + r0 * n->cfg.nin, nullptr, r);
        if (rc != SP_OK) return rc;
        MLP_HIP(hipEventRecord(n->ev[0], n->st));
        rc = fused ? fused_forward(&n, 1, n->xin, r, n->st) : forward(n, r, false);
        if (rc != SP_OK) return rc;
        MLP_HIP(hipEventRecord(n->ev[1], n->st));
        MLP_HIP(hipMemcpyAsync(y + r0 * n->cfg.nout, n->y, 4 * (size_t)r * n->cfg.nout,
                               hipMemcpyDeviceToHost, n->st));
        MLP_HIP(hipStreamSynchronize(n->st));
        float ms = 0.f;
        MLP_HIP(hipEventElapsedTime(&ms, n->ev[0], n->ev[1]));
        n->fwd_ms += ms;
    }
    MLP_HIP(hipStreamSynchronize(n->st));
    return SP_OK;
}

int sp_mlp_last_forward_ms(void* q, double* ms) {
    int rc;
    Mlp* n = live(q, rc);
    if (!n) return rc;
    if (!ms) return sp::set_error(SP_ERR_ARG, "ms is NULL");
    *ms = n->fwd_ms;
    return SP_OK;
}

int sp_mlp_ensemble_spectra(void* const* nets, int32_t members, const float* x, int64_t rows,
                            double floor, float* spec, float* spread) {
    return ensemble(nets, members, x, rows, floor, spec, spread, false);
}

int sp_mlp_ensemble_spectra_device(void* const* nets, int32_t members, const float* x, int64_t rows,
                                   double floor, float* spec, float* spread) {
    return ensemble(nets, members, x, rows, floor, spec, spread, true);
}

int sp_mlp_gradients(void* q, const float* x, const float* t, int64_t rows, float* grads, double* loss) {
    int rc;
    Mlp* n = live(q, rc);
    if (!n) return rc;
    if (rows < 1 || rows > n->bmax || !x || !t) return sp::set_error(SP_ERR_ARG, "rows must be in [1, batch_max] with x and t given");
    rc = load_rows(n, x, t, (int)rows);
    if (rc == SP_OK) rc = forward(n, (int)rows);
    if (rc == SP_OK) rc = backward(n, (int)rows, n->loss_hist);
    if (rc != SP_OK) return rc;
    float l = 0.0f;
    MLP_HIP(hipMemcpyAsync(&l, n->loss_hist, 4, hipMemcpyDeviceToHost, n->st));
    if (grads) MLP_HIP(hipMemcpyAsync(grads, n->g, 4 * (size_t)n->P, hipMemcpyDeviceToHost, n->st));
    MLP_HIP(hipStreamSynchronize(n->st));
    if (loss) *loss = l;
    return SP_OK;
}

int sp_mlp_train_step(void* q, const float* x, const float* t, int64_t rows, double lr, double* loss) {
    int rc = sp_mlp_gradients(q, x, t, rows, nullptr, loss);
    if (rc != SP_OK) return rc;
    Mlp* n = static_cast<Mlp*>(q);
    rc = adamw(n, lr);
    if (rc != SP_OK) return rc;
    MLP_HIP(hipStreamSynchronize(n->st));
    return SP_OK;
}

int sp_mlp_set_data(void* q, const float* x, const float* t, int64_t rows) {
    int rc;
    Mlp* n = live(q, rc);
    if (!n) return rc;
    if (rows < 1 || rows > 0xFFFFFFFFll || !x || !t) return sp::set_error(SP_ERR_ARG, "bad rows or NULL data");
    dfree(n->xd);
    dfree(n->td);
    n->xd = nullptr;
    n->td = nullptr;
    n->nd = 0;
    MLP_HIP(hipMalloc(&n->xd, 4 * (size_t)rows * n->cfg.nin));
    MLP_HIP(hipMalloc(&n->td, 4 * (size_t)rows * n->cfg.nout));
    MLP_HIP(hipMemcpyAsync(n->xd, x, 4 * (size_t)rows * n->cfg.nin, hipMemcpyHostToDevice, n->st));
    MLP_HIP(hipMemcpyAsync(n->td, t, 4 * (size_t)rows * n->cfg.nout, hipMemcpyHostToDevice, n->st));
    MLP_HIP(hipStreamSynchronize(n->st));
    n->nd = rows;
    return SP_OK;
}

int sp_mlp_train(void* q, int64_t nsteps, int32_t batch, double lr0, double lr_min, uint64_t seed,
                 float* losses) {
    int rc;
    Mlp* n = live(q, rc);
    if (!n) return rc;
    if (n->nd < 1) return sp::set_error(SP_ERR_STATE, "no training data (sp_mlp_set_data)");
    if (nsteps < 0 || batch < 1 || batch > n->bmax) return sp::set_error(SP_ERR_ARG, "need nsteps >= 0 and batch in [1, batch_max]");
    if (!(lr0 > 0) || !(lr_min >= 0)) return sp::set_error(SP_ERR_ARG, "learning rates must be positive");
    if (nsteps > n->loss_cap) {
        dfree(n->loss_hist);
        n->loss_hist = nullptr;
        n->loss_cap = 0;
        MLP_HIP(hipMalloc(&n->loss_hist, 4 * (size_t)nsteps));
        n->loss_cap = nsteps;
    }
    if (!n->ts) MLP_HIP(hipMalloc(&n->ts, sizeof(sp::mlp::TrainState)));
    if (nsteps > 0xFFFFFFFFll - n->step) return sp::set_error(SP_ERR_ARG, "too many steps");
    sp::mlp::TrainState hs{};
    hs.t = hs.t0 = (uint32_t)n->step;
    hs.nsteps = (uint32_t)nsteps;
    hs.lr0 = (float)lr0;
    hs.lr_min = (float)lr_min;
    hs.b1 = (float)n->cfg.beta1;
    hs.b2 = (float)n->cfg.beta2;
    n->pver += 1;
    MLP_HIP(hipMemcpyAsync(n->ts, &hs, sizeof(hs), hipMemcpyHostToDevice, n->st));
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    // One training step: every kernel reads the step from n->ts, so the same
    // launch sequence serves every step (and is captured once as a graph).
    auto one_step = [&]() -> int {
        hipLaunchKernelGGL(sp::mlp::batch_index_dev_kernel, dim3(blocks(batch)), dim3(256), 0, n->st,
                           n->bidx, batch, (uint32_t)n->nd, n->ts, k0, k1);
        hipLaunchKernelGGL(sp::mlp::gather_kernel, dim3((unsigned)batch), dim3(64), 0, n->st, n->xd,
                           n->td, n->bidx, batch, n->cfg.nin, n->cfg.nout, n->xh[0], n->t);
        MLP_HIP(hipGetLastError());
        int r = forward(n, batch);
        if (r == SP_OK) r = backward(n, batch, n->loss_hist, n->ts);
        if (r != SP_OK) return r;
        hipLaunchKernelGGL(sp::mlp::adamw_dev_kernel, dim3(blocks(n->P)), dim3(256), 0, n->st, n->p,
                           n->g, n->m, n->v, n->pb, n->P, n->ts, (float)n->cfg.eps,
                           (float)n->cfg.weight_decay);
        hipLaunchKernelGGL(sp::mlp::train_tick_kernel, dim3(1), dim3(64), 0, n->st, n->ts);
        MLP_HIP(hipGetLastError());
        return SP_OK;
    };
    const char* eager_env = std::getenv("SP_MLP_EAGER");
    const bool eager = eager_env && std::atoi(eager_env) != 0;
    if (eager || nsteps < 2) {
        for (int64_t s = 0; s < nsteps; ++s) {
            rc = one_step();
            if (rc != SP_OK) return rc;
        }
    } else {
        // the step's ~30 launches are recorded once and replayed nsteps times
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        MLP_HIP(hipStreamBeginCapture(n->st, hipStreamCaptureModeThreadLocal));
        rc = one_step();
        hipError_t e = hipStreamEndCapture(n->st, &graph);
        if (rc != SP_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        for (int64_t s = 0; s < nsteps && e == hipSuccess; ++s) e = hipGraphLaunch(exec, n->st);
        if (e == hipSuccess) e = hipStreamSynchronize(n->st);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (e != hipSuccess)
            return sp::set_error(SP_ERR_HIP, std::string("training graph failed: ") + hipGetErrorString(e));
    }
    n->step += nsteps;
    if (losses && nsteps > 0)
        MLP_HIP(hipMemcpyAsync(losses, n->loss_hist, 4 * (size_t)nsteps, hipMemcpyDeviceToHost, n->st));
    MLP_HIP(hipStreamSynchronize(n->st));
    return SP_OK;
}

}  // extern "C"
