// sp_nodes.h -- the 32 farray components of one node (8 fields and the 24
// gradients of calc_fields_gradients, mhd_data_parallel.f90:504-605) from
// the node-major 8-field array, for kernels that need more than the packed
// records hold.  Include only from translation units built with
// -ffp-contract=off: each component then rounds exactly as in the Fortran.
#pragma once

#include "sp_internal.h"

namespace sp {

// Derivative of 8 node fields along one axis at position i of an axis with
// n nodes, stride `st` (in nodes) in the node-major 8-field array.
__device__ __forceinline__ void axis_grad(const float* __restrict__ n8, long long node, int i, int n,
                                          long long st, double scale, float* out, int comp0) {
    if (n <= 1) {
#pragma unroll
        for (int c = 0; c < 8; ++c) out[comp0 + 3 * c] = 0.0f;
        return;
    }
    if (i >= 1 && i <= n - 2) {
        const float* a = n8 + (node + st) * 8;
        const float* b = n8 + (node - st) * 8;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const float d = a[c] - b[c];
            out[comp0 + 3 * c] = (float)((double)d * scale);
        }
    } else if (i == 0) {
        const float* f0 = n8 + node * 8;
        const float* f1 = n8 + (node + st) * 8;
        const float* f2 = n8 + (node + 2 * st) * 8;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const float d = (-3.0f * f0[c] + 4.0f * f1[c]) - f2[c];
            out[comp0 + 3 * c] = (float)((double)d * scale);
        }
    } else {
        const float* f0 = n8 + node * 8;
        const float* f1 = n8 + (node - st) * 8;
        const float* f2 = n8 + (node - 2 * st) * 8;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const float d = (3.0f * f0[c] - 4.0f * f1[c]) + f2[c];
            out[comp0 + 3 * c] = (float)((double)d * scale);
        }
    }
}

// The 32 farray components of one node (8 fields + 24 gradients).
__device__ __forceinline__ void node_components(const DevParams& prm, double idxh, double idyh,
                                                double idzh, const float* __restrict__ n8,
                                                long long node, float out[32]) {
    const int i = (int)(node % prm.nxg);
    const int j = (int)((node / prm.nxg) % prm.nyg);
    const int k = (int)(node / ((long long)prm.nxg * prm.nyg));
#pragma unroll
    for (int c = 0; c < 8; ++c) out[c] = n8[node * 8 + c];
    axis_grad(n8, node, i, prm.nxg, 1, idxh, out, 8);
    axis_grad(n8, node, j, prm.nyg, prm.nxg, idyh, out, 9);
    axis_grad(n8, node, k, prm.nzg, (long long)prm.nxg * prm.nyg, idzh, out, 10);
}

}  // namespace sp
