// sp_sort.hip -- spatial ordering of the particle arrays.
//
// The mover hands out particles in array order, so the particles in flight at
// any moment are a window of consecutive indices.  Ordering the arrays by a
// Morton key of coarse grid blocks makes that window a compact region of the
// grid whose field records stay resident in L2/MALL, instead of random cells
// all over the snapshot.  The order never changes a result: every particle's
// increments are keyed on its own identity.
//
// Stable LSD radix sort of (key, index) pairs, 4-bit digits: per-tile digit
// histograms, one exclusive scan of the digit-major tile table, and a stable
// scatter that ranks equal digits inside each 64-lane wave with ballots.  The
// resulting permutation gathers all particle arrays into the spare buffer set.
#include "sp_device.h"
#include "sp_particles.h"

namespace sp {

namespace {

constexpr int kBlock = 256;
constexpr int kTile = kBlock * kTileItems;
constexpr int kDigits = 16;

__device__ __forceinline__ uint32_t spread2(uint32_t v) {  // 16 bits -> even bits
    v &= 0xFFFFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

__device__ __forceinline__ uint32_t spread3(uint32_t v) {  // 10 bits -> every third bit
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__device__ __forceinline__ uint32_t block_of(double pos, double lo, double inv_d, int ncell, int shift) {
    double c = floor((pos - lo) * inv_d);
    c = fmin(fmax(c, 0.0), (double)(ncell - 1));
    return (uint32_t)c >> shift;
}

__global__ void __launch_bounds__(kBlock) key_kernel(Ptl P, long long n, SortArgs sa, uint32_t* key,
                                                     uint32_t* idx) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t bx = block_of(P.x[i], sa.xmin, sa.idx, sa.nx, sa.shift);
    const uint32_t by = block_of(P.y[i], sa.ymin, sa.idy, sa.ny, sa.shift);
    uint32_t k;
    if (sa.ndim == 3) {
        const uint32_t bz = block_of(P.z[i], sa.zmin, sa.idz, sa.nz, sa.shift);
        k = spread3(bx) | (spread3(by) << 1) | (spread3(bz) << 2);
    } else {
        k = spread2(bx) | (spread2(by) << 1);
    }
    if (sa.done_bit >= 0 && !(P.rdt[i] > 0.0)) k |= 1u << sa.done_bit;
    key[i] = k;
    idx[i] = (uint32_t)i;
}

__global__ void __launch_bounds__(kBlock) digit_count_kernel(const uint32_t* key, long long n, int shift,
                                                             long long ntiles, unsigned* hist) {
    __shared__ unsigned h[kDigits];
    if (threadIdx.x < kDigits) h[threadIdx.x] = 0;
    __syncthreads();
    const long long t0 = (long long)blockIdx.x * kTile;
#pragma unroll
    for (int it = 0; it < kTileItems; ++it) {
        const long long i = t0 + (long long)it * kBlock + threadIdx.x;
        if (i < n) atomicAdd(&h[(key[i] >> shift) & (kDigits - 1)], 1u);
    }
    __syncthreads();
    if (threadIdx.x < kDigits) hist[(long long)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

__global__ void __launch_bounds__(kBlock) digit_scatter_kernel(const uint32_t* kin, const uint32_t* iin,
                                                               uint32_t* kout, uint32_t* iout,
                                                               long long n, int shift,
                                                               long long ntiles,
                                                               const unsigned long long* off) {
    __shared__ unsigned wcnt[kBlock / kWave][kDigits];
    __shared__ unsigned long long run[kDigits];
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    if (threadIdx.x < kDigits) run[threadIdx.x] = off[(long long)threadIdx.x * ntiles + blockIdx.x];
    __syncthreads();
    const long long t0 = (long long)blockIdx.x * kTile;
    const unsigned long long lt = (1ull << lane) - 1ull;
#pragma unroll 1
    for (int it = 0; it < kTileItems; ++it) {
        const long long i = t0 + (long long)it * kBlock + threadIdx.x;
        const bool valid = i < n;
        const uint32_t k = valid ? kin[i] : 0u;
        const int d = valid ? (int)((k >> shift) & (kDigits - 1)) : kDigits;
        unsigned rank = 0;
#pragma unroll
        for (int v = 0; v < kDigits; ++v) {
            const unsigned long long b = __ballot(d == v);
            if (d == v) rank = (unsigned)__popcll(b & lt);
            if (lane == 0) wcnt[wv][v] = (unsigned)__popcll(b);
        }
        __syncthreads();
        if (valid) {
            unsigned long long o = run[d] + rank;
            for (int q = 0; q < wv; ++q) o += wcnt[q][d];
            kout[o] = k;
            iout[o] = iin[i];
        }
        __syncthreads();
        if (threadIdx.x < kDigits) {
            unsigned tot = 0;
            for (int q = 0; q < kBlock / kWave; ++q) tot += wcnt[q][threadIdx.x];
            run[threadIdx.x] += tot;
        }
        __syncthreads();
    }
}

// D[j] = S[perm[j]] for j < n (perm == nullptr: the identity).
__global__ void __launch_bounds__(kBlock) permute_kernel(Ptl S, Ptl D, long long n, const uint32_t* perm) {
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const long long i = perm ? (long long)perm[j] : j;
    store_row<true>(D, j, load_row(S, i));
}

}  // namespace

hipError_t launch_spatial_sort(const Ptl& S, const Ptl& D, long long n, const SortArgs& sa,
                               uint32_t* keyA, uint32_t* idxA, uint32_t* keyB, uint32_t* idxB,
                               unsigned* hist, unsigned long long* off, unsigned long long* totals,
                               bool copy_back, hipStream_t s) {
    if (n <= 1) return hipSuccess;
    const unsigned grid = (unsigned)((n + kBlock - 1) / kBlock);
    const long long ntiles = num_tiles(n);
    hipLaunchKernelGGL(key_kernel, dim3(grid), dim3(kBlock), 0, s, S, n, sa, keyA, idxA);
    for (int pass = 0; pass < sa.passes; ++pass) {
        const int shift = 4 * pass;
        hipLaunchKernelGGL(digit_count_kernel, dim3((unsigned)ntiles), dim3(kBlock), 0, s, keyA, n,
                           shift, ntiles, hist);
        hipError_t e = launch_tile_scan(hist, nullptr, kDigits * ntiles, off, totals, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(digit_scatter_kernel, dim3((unsigned)ntiles), dim3(kBlock), 0, s, keyA,
                           idxA, keyB, idxB, n, shift, ntiles, off);
        uint32_t* t = keyA; keyA = keyB; keyB = t;
        t = idxA; idxA = idxB; idxB = t;
    }
    hipLaunchKernelGGL(permute_kernel, dim3(grid), dim3(kBlock), 0, s, S, D, n, idxA);
    if (copy_back) {
        // a sorted prefix goes back over the prefix of S; the rest of S stays
        hipLaunchKernelGGL(permute_kernel, dim3(grid), dim3(kBlock), 0, s, D, S, n,
                           (const uint32_t*)nullptr);
    }
    return hipGetLastError();
}

}  // namespace sp
