// Block-wise max |D| over the (group, group) AO blocks of many densities
// (the density bound of the prescreen, cuobc/scf/hf.py:104-107).
#include <hip/hip_runtime.h>

#define BQC_DMC_MAXG 256

#define BQC_DMC_ROWS 8

__device__ __forceinline__ int group_of(const int *L, int G, int r) {
    int lo = 0, hi = G;  // L[lo] <= r < L[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (L[mid] <= r) lo = mid;
        else hi = mid;
    }
    return lo;
}

// mats: nmat matrices; for matrix m, rows/columns of group g are
// [loc[off[m] + g], loc[off[m] + g + 1]); ng[m] groups; out[m] holds
// ng[m] x ng[m] values at out + outoff[m] and must start at zero.
// Workgroup (x, m) scans rows [x*R, x*R + R) of matrix m; per row, lanes
// stride over the columns of each column group and the block max goes to
// LDS, flushed with one global atomicMax per (row group, column group) when
// the row group changes or the tile ends.
__global__ void __launch_bounds__(256)
    dm_cond_kernel(const double *const *mats, const int *nao, const int *loc, const int *off,
                   const int *ng, const int *outoff, double *out) {
    const int m = blockIdx.y;
    const int n = nao[m];
    const int r0 = blockIdx.x * BQC_DMC_ROWS;
    if (r0 >= n) return;
    const int G = ng[m];
    const int *L = loc + off[m];
    const double *D = mats[m];
    unsigned long long *O = (unsigned long long *)(out + outoff[m]);
    __shared__ unsigned long long colmax[BQC_DMC_MAXG];
    for (int h = threadIdx.x; h < G; h += blockDim.x) colmax[h] = 0ull;
    __syncthreads();
    const int r1 = min(r0 + BQC_DMC_ROWS, n);
    int g = group_of(L, G, r0);
    for (int r = r0; r < r1; ++r) {
        for (int h = 0; h < G; ++h) {
            double mx = 0.0;
            for (int c = L[h] + threadIdx.x; c < L[h + 1]; c += blockDim.x)
                mx = fmax(mx, fabs(D[(size_t)r * n + c]));
            for (int d = 32; d >= 1; d >>= 1) mx = fmax(mx, __shfl_xor(mx, d, 64));
            // non-negative doubles order like their bit patterns
            if ((threadIdx.x & 63) == 0 && mx > 0.0)
                atomicMax(&colmax[h], (unsigned long long)__double_as_longlong(mx));
        }
        const bool flush = r + 1 == r1 || r + 1 >= L[g + 1];
        if (flush) {
            __syncthreads();
            for (int h = threadIdx.x; h < G; h += blockDim.x) {
                if (colmax[h]) atomicMax(&O[g * G + h], colmax[h]);
                colmax[h] = 0ull;
            }
            __syncthreads();
            if (r + 1 < r1) g = group_of(L, G, r + 1);
        }
    }
}

extern "C" int BQCdm_cond(const double *const *mats, const int *nao, const int *loc,
                          const int *off, const int *ng, const int *outoff, double *out, int nmat,
                          int gmax, int naomax, hipStream_t stream) {
    if (gmax > BQC_DMC_MAXG || nmat <= 0 || naomax <= 0) return 1;
    const int nblk = (naomax + BQC_DMC_ROWS - 1) / BQC_DMC_ROWS;
    hipLaunchKernelGGL(dm_cond_kernel, dim3(nblk, nmat), dim3(256), 0, stream, mats, nao, loc,
                       off, ng, outoff, out);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}
