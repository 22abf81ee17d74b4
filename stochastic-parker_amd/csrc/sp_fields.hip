// sp_fields.hip -- field preparation on the device: the 24 gradient
// components of calc_fields_gradients (mhd_data_parallel.f90:504-605) and the
// packing of each node into the record the mover gathers.
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

__global__ void __launch_bounds__(256) gradient_max_kernel(DevParams prm, double idxh, double idyh,
                                                           double idzh, const float* __restrict__ n8,
                                                           unsigned* maxbits) {
    __shared__ unsigned red[256];
    const long long total = (long long)prm.nxg * prm.nyg * prm.nzg;
    const long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float mx = 0.0f;
    if (node < total) {
        float out[32];
        node_components(prm, idxh, idyh, idzh, n8, node, out);
#pragma unroll
        for (int s = R3_DVXDX; s <= R3_DBDZ; ++s) mx = fmaxf(mx, fabsf(out[rec3_src(s)]));
    }
    red[threadIdx.x] = __float_as_uint(mx);   // non-negative floats order as integers
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + st]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(maxbits, red[0]);
}

__global__ void __launch_bounds__(256) pack_half_kernel(DevParams prm, double idxh, double idyh,
                                                        double idzh, const float* __restrict__ n8,
                                                        float scale, float* __restrict__ rec,
                                                        float* __restrict__ fa32) {
    const long long total = (long long)prm.nxg * prm.nyg * prm.nzg;
    const long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= total) return;
    float out[32];
    node_components(prm, idxh, idyh, idzh, n8, node, out);
    uint32_t wd[kRec3DH];
#pragma unroll
    for (int s = 0; s < 6; ++s) wd[s] = __float_as_uint(out[rec3_src(s)]);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int s0 = R3_DVXDX + 2 * q, s1 = s0 + 1;
        const _Float16 h0 = (_Float16)(out[rec3_src(s0)] * scale);
        const _Float16 h1 = s1 <= R3_DBDZ ? (_Float16)(out[rec3_src(s1)] * scale) : (_Float16)0.0f;
        wd[6 + q] = (uint32_t)__builtin_bit_cast(uint16_t, h0) |
                    ((uint32_t)__builtin_bit_cast(uint16_t, h1) << 16);
    }
    wd[14] = wd[15] = 0u;
    const int i = (int)(node % prm.nxg);
    const int j = (int)((node / prm.nxg) % prm.nyg);
    const int k = (int)(node / ((long long)prm.nxg * prm.nyg));
    uint4* r4 = reinterpret_cast<uint4*>(rec + brick_record(i, j, k, prm.bnx, prm.bny) * kRec3DH);
#pragma unroll
    for (int u = 0; u < kRec3DH / 4; ++u) r4[u] = make_uint4(wd[4 * u], wd[4 * u + 1], wd[4 * u + 2], wd[4 * u + 3]);
    if (fa32) {
#pragma unroll
        for (int c = 0; c < 32; ++c) fa32[node * 32 + c] = out[c];
    }
}

}  // namespace

hipError_t launch_gradient_max(const DevParams& prm, const float* nodes8, unsigned* maxbits,
                               hipStream_t s) {
    const long long total = (long long)prm.nxg * prm.nyg * prm.nzg;
    const long long grid = (total + 255) / 256;
    hipLaunchKernelGGL(gradient_max_kernel, dim3((unsigned)grid), dim3(256), 0, s, prm, 0.5 / prm.dx,
                       0.5 / prm.dy, 0.5 / prm.dz, nodes8, maxbits);
    return hipGetLastError();
}

hipError_t launch_pack_half(const DevParams& prm, const float* nodes8, float scale, float* rec,
                            float* farray32, hipStream_t s) {
    const long long total = (long long)prm.nxg * prm.nyg * prm.nzg;
    const long long grid = (total + 255) / 256;
    hipLaunchKernelGGL(pack_half_kernel, dim3((unsigned)grid), dim3(256), 0, s, prm, 0.5 / prm.dx,
                       0.5 / prm.dy, 0.5 / prm.dz, nodes8, scale, rec, farray32);
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
