// Block-wise max |D| over the (group, group) AO blocks of many densities
// (the density bound of the prescreen, cuobc/scf/hf.py:104-107).  One
// workgroup per (matrix, block row group); lanes stride over the rows of the
// group and all columns, keeping one running max per column group in LDS.
#include <hip/hip_runtime.h>

#define BQC_DMC_MAXG 256

// mats: nmat matrices; for matrix m, rows/columns of group g are
// [loc[off[m] + g], loc[off[m] + g + 1]); ng[m] groups; out[m] holds
// ng[m] x ng[m] values at out + outoff[m].
__global__ void dm_cond_kernel(const double *const *mats, const int *nao, const int *loc,
                               const int *off, const int *ng, const int *outoff, double *out) {
    const int m = blockIdx.y, g = blockIdx.x;
    const int G = ng[m];
    if (g >= G) return;
    __shared__ double colmax[BQC_DMC_MAXG];
    for (int h = threadIdx.x; h < G; h += blockDim.x) colmax[h] = 0.0;
    __syncthreads();
    const double *D = mats[m];
    const int n = nao[m];
    const int *L = loc + off[m];
    const int r0 = L[g], r1 = L[g + 1];
    for (int h = 0; h < G; ++h) {
        const int c0 = L[h], c1 = L[h + 1];
        const int w = c1 - c0;
        double mx = 0.0;
        for (long long e = threadIdx.x; e < (long long)(r1 - r0) * w; e += blockDim.x) {
            const int r = r0 + (int)(e / w), c = c0 + (int)(e % w);
            mx = fmax(mx, fabs(D[(size_t)r * n + c]));
        }
        // wave max then LDS max
        for (int d = 32; d >= 1; d >>= 1) mx = fmax(mx, __shfl_xor(mx, d, 64));
        if ((threadIdx.x & 63) == 0) {
            unsigned long long *p = (unsigned long long *)&colmax[h];
            // non-negative doubles order like their bit patterns
            atomicMax(p, (unsigned long long)__double_as_longlong(mx));
        }
    }
    __syncthreads();
    for (int h = threadIdx.x; h < G; h += blockDim.x) out[outoff[m] + g * G + h] = colmax[h];
}

extern "C" int BQCdm_cond(const double *const *mats, const int *nao, const int *loc,
                          const int *off, const int *ng, const int *outoff, double *out, int nmat,
                          int gmax, hipStream_t stream) {
    if (gmax > BQC_DMC_MAXG || nmat <= 0) return 1;
    hipLaunchKernelGGL(dm_cond_kernel, dim3(gmax, nmat), dim3(256), 0, stream, mats, nao, loc,
                       off, ng, outoff, out);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}
