// refine.hip -- FP64 iterative refinement around the factorization, for the reduced-precision
// update forms: x0 = solve(b), x_{k+1} = x_k + solve(b - A x_k), stopped by the scaled
// backward error ||A x - b|| / (||A||_F ||x|| + ||b||) of sparse_backward_error
// (sparse/src/validate.rs:209-255), evaluated on the device from the full-symmetric CSC.
// The reference's solve has no refinement (solve.rs:57-63).
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <cmath>

#include "rv_numeric.h"

namespace rv {

namespace {

constexpr int RB = 256;  // threads per block of the reductions

// r = b - A x.  Both triangles are stored, so (A x)_j is the dot product of column j with x.
__global__ __launch_bounds__(RB) void k_residual(long long n, const long long *__restrict__ cp,
                                                 const long long *__restrict__ ri,
                                                 const double *__restrict__ val,
                                                 const double *__restrict__ x,
                                                 const double *__restrict__ b, double *__restrict__ r) {
    const long long j = (long long)blockIdx.x * RB + threadIdx.x;
    if (j >= n) return;
    double s = 0.0;
    for (long long p = cp[j]; p < cp[j + 1]; ++p) s += val[p] * x[ri[p]];
    r[j] = b[j] - s;
}

// per-block sum of squares in a fixed tree order
__global__ __launch_bounds__(RB) void k_sumsq(long long n, const double *__restrict__ a,
                                              double *__restrict__ part) {
    __shared__ double sh[RB];
    double s = 0.0;
    for (long long i = (long long)blockIdx.x * RB + threadIdx.x; i < n; i += (long long)gridDim.x * RB)
        s += a[i] * a[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = RB / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(RB) void k_sum_parts(int np, const double *__restrict__ part,
                                                  double *__restrict__ out) {
    __shared__ double sh[RB];
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += RB) s += part[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = RB / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

__global__ __launch_bounds__(RB) void k_add(long long n, double *__restrict__ x, const double *__restrict__ d) {
    const long long i = (long long)blockIdx.x * RB + threadIdx.x;
    if (i < n) x[i] += d[i];
}

constexpr int NPART = 1024;

// sum of squares of a[0..n) into *d_out (device), deterministic
void sumsq(hipStream_t st, long long n, const double *a, double *d_part, double *d_out) {
    const int nb = (int)std::min<long long>(NPART, std::max<long long>(1, (n + RB - 1) / RB));
    hipLaunchKernelGGL(k_sumsq, dim3(nb), dim3(RB), 0, st, n, a, d_part);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(RB), 0, st, nb, (const double *)d_part, d_out);
}

}  // namespace

RefineResult solve_refine(NumericGPU &gpu, hipStream_t st, i64 n, i64 nnz, const i64 *d_cp, const i64 *d_ri,
                          const double *d_val, const double *d_scal, const double *d_b, double *d_x,
                          int max_iter, double tol, DevBuf &work) {
    RefineResult res;
    // work: r (n), partials (NPART), four sums
    double *r = (double *)work.get(((size_t)n + NPART + 8) * sizeof(double));
    double *part = r + n;
    double *sums = part + NPART;  // ||A||_F^2, ||b||^2, ||r||^2, ||x||^2
    const dim3 g((unsigned)((n + RB - 1) / RB)), b(RB);
    sumsq(st, nnz, d_val, part, sums + 0);
    sumsq(st, n, d_b, part, sums + 1);
    RV_HIP(hipMemcpyAsync(d_x, d_b, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, st));
    gpu.solve(d_x, n, 1, d_scal);
    double h[4];
    for (int it = 0;; ++it) {
        hipLaunchKernelGGL(k_residual, g, b, 0, st, (long long)n, (const long long *)d_cp,
                           (const long long *)d_ri, d_val, (const double *)d_x, d_b, r);
        sumsq(st, n, r, part, sums + 2);
        sumsq(st, n, d_x, part, sums + 3);
        RV_HIP(hipMemcpyAsync(h, sums, 4 * sizeof(double), hipMemcpyDeviceToHost, st));
        RV_HIP(hipStreamSynchronize(st));
        const double denom = std::sqrt(h[0]) * std::sqrt(h[3]) + std::sqrt(h[1]);
        res.backward_error = denom == 0.0 ? 0.0 : std::sqrt(h[2]) / denom;
        res.iters = it;
        if (!(res.backward_error > tol) || it >= max_iter) break;
        gpu.solve(r, n, 1, d_scal);
        hipLaunchKernelGGL(k_add, g, b, 0, st, (long long)n, d_x, (const double *)r);
    }
    return res;
}

}  // namespace rv
