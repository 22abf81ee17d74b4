// sp_fields.hip -- field preparation on the device: the 24 gradient
// components of calc_fields_gradients (mhd_data_parallel.f90:504-605) and the
// packing of each node into the record the mover gathers (node records for
// the FP64 mover and the momentum-diffusion steps, the compact layouts of
// the FP32 movers).
//
// Built with -ffp-contract=off: the FP32 differences, the FP64 scaling and the
// FP32 store round exactly as in the Fortran, so the gradients are bit-exact
// with the CPU restatement.
#include "sp_nodes.h"

namespace sp {

namespace {

// One node record: the first NC slots of the record layout, zero padding
// up to W floats, stored as W/4 dwordx4.
template <int W, int NC, int NDIM>
__device__ __forceinline__ void pack_record(const float* out, float* rec) {
    float4* r4 = reinterpret_cast<float4*>(rec);
    float v[W];
#pragma unroll
    for (int s = 0; s < W; ++s) v[s] = s < NC ? out[NDIM == 2 ? rec2_src(s) : rec3_src(s)] : 0.0f;
#pragma unroll
    for (int u = 0; u < W / 4; ++u) r4[u] = make_float4(v[4 * u], v[4 * u + 1], v[4 * u + 2], v[4 * u + 3]);
}

__global__ void __launch_bounds__(256) gradients_kernel(DevParams prm, double idxh, double idyh,
                                                        double idzh, const float* __restrict__ n8,
                                                        float* __restrict__ rec,
                                                        float* __restrict__ fa32) {
    const long long total = (long long)prm.nxg * prm.nyg * prm.nzg;
    const long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= total) return;
    const int i = (int)(node % prm.nxg);
    const int j = (int)((node / prm.nxg) % prm.nyg);
    const int k = (int)(node / ((long long)prm.nxg * prm.nyg));
    float out[32];
#pragma unroll
    for (int c = 0; c < 8; ++c) out[c] = n8[node * 8 + c];
    axis_grad(n8, node, i, prm.nxg, 1, idxh, out, 8);
    axis_grad(n8, node, j, prm.nyg, prm.nxg, idyh, out, 9);
    axis_grad(n8, node, k, prm.nzg, (long long)prm.nxg * prm.nyg, idzh, out, 10);
    if (prm.ndim == 2) {
        if (prm.ext_rec) pack_record<kRec2DX, 18, 2>(out, rec + node * kRec2DX);
        else pack_record<kRec2D, 15, 2>(out, rec + node * kRec2D);
    } else {
        const long long r = brick_record(i, j, k, prm.bnx, prm.bny);
        if (prm.ext_rec) pack_record<kRec3DX, 28, 3>(out, rec + r * kRec3DX);
        else pack_record<kRec3D, 21, 3>(out, rec + r * kRec3D);
    }
    if (fa32) {
#pragma unroll
        for (int c = 0; c < 32; ++c) fa32[node * 32 + c] = out[c];
    }
}

// The FP32-mover slots of one node (pair2_src / node3_src order) from its
// 32 farray components.
template <int NDIM>
__device__ __forceinline__ void fast_values(const float* out, float* v) {
#pragma unroll
    for (int s = 0; s < (NDIM == 2 ? Pair2Shape::N : 19); ++s) {
        const int src = NDIM == 2 ? pair2_src(s) : node3_src(s);
        if (src >= 0) v[s] = out[src];
        else v[s] = NDIM == 2 ? out[8] + out[12] : (out[8] + out[12]) + out[16];
    }
}

// maxbits[0] = max |gradient slot| (divV included), maxbits[1] = max |V|
template <int NDIM>
__global__ void __launch_bounds__(256) value_max_kernel(DevParams prm, double idxh, double idyh,
                                                        double idzh, const float* __restrict__ n8,
                                                        unsigned* maxbits) {
    constexpr int NF = NDIM == 2 ? Pair2Shape::NF : 6;
    constexpr int N = NDIM == 2 ? Pair2Shape::N : 19;
    __shared__ unsigned red[2][256];
    const long long total = (long long)prm.nxg * prm.nyg * prm.nzg;
    const long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float mg = 0.0f, mv = 0.0f;
    if (node < total) {
        float out[32], v[N];
        node_components(prm, idxh, idyh, idzh, n8, node, out);
        fast_values<NDIM>(out, v);
#pragma unroll
        for (int s = NF; s < N; ++s) mg = fmaxf(mg, fabsf(v[s]));
#pragma unroll
        for (int s = 0; s < NDIM; ++s) mv = fmaxf(mv, fabsf(v[s]));
    }
    red[0][threadIdx.x] = __float_as_uint(mg);   // non-negative floats order as integers
    red[1][threadIdx.x] = __float_as_uint(mv);
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            red[0][threadIdx.x] = max(red[0][threadIdx.x], red[0][threadIdx.x + st]);
            red[1][threadIdx.x] = max(red[1][threadIdx.x], red[1][threadIdx.x + st]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicMax(maxbits, red[0][0]);
        atomicMax(maxbits + 1, red[1][0]);
    }
}

__device__ __forceinline__ uint32_t pack_halves(float a, float b) {
    return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)a) |
           ((uint32_t)__builtin_bit_cast(uint16_t, (_Float16)b) << 16);
}

// One thread per node, the record layouts of the FP32 movers (kCell2D,
// Pair2Shape, kRec3F, kRec3H).  2D FP32: the coefficient record of the cell
// whose corner 0 is the node.  2D FP16 gradients: the x-pair record of node
// (i, j) from the node and its +x neighbour (zeros past the last node of the
// row, which no cell reads), node-major.  3D: the node's own record, bricked.
template <int NDIM, bool HALF>
__global__ void __launch_bounds__(256) pack_fast_kernel(DevParams prm, double idxh, double idyh,
                                                        double idzh, const float* __restrict__ n8,
                                                        float gscale, float vscale,
                                                        float* __restrict__ rec,
                                                        float* __restrict__ fa32) {
    using S = Pair2Shape;
    constexpr int W = NDIM == 2 ? (HALF ? S::W16 : kCell2D) : (HALF ? kRec3H : kRec3F);
    constexpr int N = NDIM == 2 ? S::N : 19;
    const long long total = (long long)prm.nxg * prm.nyg * prm.nzg;
    const long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= total) return;
    const int i = (int)(node % prm.nxg);
    const int j = (int)((node / prm.nxg) % prm.nyg);
    const int k = (int)(node / ((long long)prm.nxg * prm.nyg));
    float out[32], v[2][N];
    node_components(prm, idxh, idyh, idzh, n8, node, out);
    if (fa32) {
#pragma unroll
        for (int c = 0; c < 32; ++c) fa32[node * 32 + c] = out[c];
    }
    fast_values<NDIM>(out, v[0]);
    uint32_t wd[W];
#pragma unroll
    for (int w = 0; w < W; ++w) wd[w] = 0u;
    if constexpr (NDIM == 2 && !HALF) {
        // bilinear coefficients of cell (i, j); zeros for the last row and
        // column of nodes, which start no cell
        if (i + 1 < prm.nxg && j + 1 < prm.nyg) {
            float c[3][N];
            const long long nb[3] = {node + 1, node + prm.nxg, node + prm.nxg + 1};
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                node_components(prm, idxh, idyh, idzh, n8, nb[q], out);
                fast_values<NDIM>(out, c[q]);
            }
#pragma unroll
            for (int s = 0; s < N; ++s) {
                const float v00 = v[0][s], v10 = c[0][s], v01 = c[1][s], v11 = c[2][s];
                wd[s] = __float_as_uint(v00);
                wd[N + s] = __float_as_uint(v10 - v00);
                wd[2 * N + s] = __float_as_uint(v01 - v00);
                wd[3 * N + s] = __float_as_uint((v11 - v10) - (v01 - v00));
            }
        }
    } else if constexpr (NDIM == 2) {
        if (i + 1 < prm.nxg) {
            node_components(prm, idxh, idyh, idzh, n8, node + 1, out);
            fast_values<NDIM>(out, v[1]);
        } else {
#pragma unroll
            for (int s = 0; s < N; ++s) v[1][s] = 0.0f;
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
            for (int s = 0; s < S::NF; ++s) wd[h * S::NF + s] = __float_as_uint(v[h][s]);
#pragma unroll
            for (int q = 0; q < S::NGW; ++q) {
                const int s0 = S::NF + 2 * q, s1 = s0 + 1;
                wd[2 * S::NF + h * S::NGW + q] =
                    pack_halves(v[h][s0] * gscale, s1 < N ? v[h][s1] * gscale : 0.0f);
            }
        }
    } else if (!HALF) {
#pragma unroll
        for (int s = 0; s < N; ++s) wd[s] = __float_as_uint(v[0][s]);
    } else {
        // B as floats; halves: V (velocity scale), then divV and the gradients
#pragma unroll
        for (int s = 0; s < 3; ++s) wd[s] = __float_as_uint(v[0][3 + s]);
        float hv[16];
#pragma unroll
        for (int s = 0; s < 3; ++s) hv[s] = v[0][s] * vscale;
#pragma unroll
        for (int s = 6; s < N; ++s) hv[s - 3] = v[0][s] * gscale;
#pragma unroll
        for (int q = 0; q < 8; ++q) wd[3 + q] = pack_halves(hv[2 * q], hv[2 * q + 1]);
    }
    const long long r = NDIM == 3 ? brick_record(i, j, k, prm.bnx, prm.bny) : node;
    uint4* r4 = reinterpret_cast<uint4*>(rec + r * W);
#pragma unroll
    for (int u = 0; u < W / 4; ++u) r4[u] = make_uint4(wd[4 * u], wd[4 * u + 1], wd[4 * u + 2], wd[4 * u + 3]);
}

}  // namespace

hipError_t launch_value_max(const DevParams& prm, const float* nodes8, unsigned* maxbits,
                            hipStream_t s) {
    const long long total = (long long)prm.nxg * prm.nyg * prm.nzg;
    const long long grid = (total + 255) / 256;
    auto k = prm.ndim == 2 ? value_max_kernel<2> : value_max_kernel<3>;
    hipLaunchKernelGGL(k, dim3((unsigned)grid), dim3(256), 0, s, prm, 0.5 / prm.dx, 0.5 / prm.dy,
                       0.5 / prm.dz, nodes8, maxbits);
    return hipGetLastError();
}

hipError_t launch_pack_fast(const DevParams& prm, const float* nodes8, bool half, float gscale,
                            float vscale, float* rec, float* farray32, hipStream_t s) {
    const long long total = (long long)prm.nxg * prm.nyg * prm.nzg;
    const long long grid = (total + 255) / 256;
    auto k = prm.ndim == 2 ? (half ? pack_fast_kernel<2, true> : pack_fast_kernel<2, false>)
                           : (half ? pack_fast_kernel<3, true> : pack_fast_kernel<3, false>);
    hipLaunchKernelGGL(k, dim3((unsigned)grid), dim3(256), 0, s, prm, 0.5 / prm.dx, 0.5 / prm.dy,
                       0.5 / prm.dz, nodes8, gscale, vscale, rec, farray32);
    return hipGetLastError();
}

hipError_t launch_gradients(const DevParams& prm, const float* nodes8, float* rec, float* farray32,
                            hipStream_t s) {
    const long long total = (long long)prm.nxg * prm.nyg * prm.nzg;
    const int block = 256;
    const long long grid = (total + block - 1) / block;
    const double idxh = 0.5 / prm.dx, idyh = 0.5 / prm.dy, idzh = 0.5 / prm.dz;
    hipLaunchKernelGGL(gradients_kernel, dim3((unsigned)grid), dim3(block), 0, s, prm, idxh, idyh,
                       idzh, nodes8, rec, farray32);
    return hipGetLastError();
}

}  // namespace sp
