// FP64 kernels of the L-BFGS method (Nocedal & Wright, Numerical
// Optimization, 2nd ed., Algorithm 7.4, the two-loop recursion).  The
// reference runs the method on the host through scipy's L-BFGS-B
// (fwi/optim_lbfgsb.py:409-434) or jaxopt.LBFGS (fwi/optimizer.py:110-117);
// here the vectors stay on the device in float64.
//
// The recursion is a chain of 2m dependent dot products over vectors of
// n = 76,000 (Vs only) to a few 10^5 entries, so one workgroup of 1024
// threads runs it: each thread owns the entries i = tid (mod 1024) of every
// vector, partial sums go through a fixed LDS tree, and the result is
// bitwise reproducible.  No inter-workgroup synchronisation is needed.
#pragma once
#include <hip/hip_runtime.h>

namespace stfwi {

#define LB_THREADS 1024
#define LB_MAX_M 64

// sum of one double per thread over the workgroup; every thread gets the
// result.  `red` is LB_THREADS doubles of LDS.
__device__ __forceinline__ double lb_block_sum(double v, double *red)
{
    const int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
#pragma unroll
    for (int h = LB_THREADS / 2; h > 0; h >>= 1) {
        if (t < h) red[t] += red[t + h];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

struct LbHist {
    int slot[LB_MAX_M];     // history rows, oldest first
    double rho[LB_MAX_M];   // 1 / (y_i . s_i)
};

// d = -H g with H the L-BFGS inverse-Hessian approximation from the m
// stored pairs and H0 = gamma I; also *gd = g . d
__global__ __launch_bounds__(LB_THREADS) void k_lbfgs_dir(const double *__restrict__ g, const double *__restrict__ S,
                                                          const double *__restrict__ Y, LbHist h, int m, long long n,
                                                          long long ld, double gamma, double *__restrict__ d,
                                                          double *gd)
{
    __shared__ double red[LB_THREADS];
    __shared__ double alpha[LB_MAX_M];
    const long long t0 = threadIdx.x;
    for (long long i = t0; i < n; i += LB_THREADS) d[i] = g[i];
    // first loop: newest to oldest
    for (int k = m - 1; k >= 0; --k) {
        const double *s = S + (long long)h.slot[k] * ld, *y = Y + (long long)h.slot[k] * ld;
        double p = 0.0;
        for (long long i = t0; i < n; i += LB_THREADS) p = fma(s[i], d[i], p);
        const double a = h.rho[k] * lb_block_sum(p, red);
        if (threadIdx.x == 0) alpha[k] = a;
        for (long long i = t0; i < n; i += LB_THREADS) d[i] = fma(-a, y[i], d[i]);
    }
    for (long long i = t0; i < n; i += LB_THREADS) d[i] *= gamma;
    __syncthreads();
    // second loop: oldest to newest
    for (int k = 0; k < m; ++k) {
        const double *s = S + (long long)h.slot[k] * ld, *y = Y + (long long)h.slot[k] * ld;
        double p = 0.0;
        for (long long i = t0; i < n; i += LB_THREADS) p = fma(y[i], d[i], p);
        const double b = h.rho[k] * lb_block_sum(p, red);
        const double c = alpha[k] - b;
        for (long long i = t0; i < n; i += LB_THREADS) d[i] = fma(c, s[i], d[i]);
    }
    double p = 0.0;
    for (long long i = t0; i < n; i += LB_THREADS) {
        d[i] = -d[i];
        p = fma(g[i], d[i], p);
    }
    const double r = lb_block_sum(p, red);
    if (threadIdx.x == 0 && gd) *gd = r;
}

// s = x_new - x_old, y = g_new - g_old; dots = (y.s, y.y, s.s)
__global__ __launch_bounds__(LB_THREADS) void k_lbfgs_pair(const double *__restrict__ xn, const double *__restrict__ xo,
                                                           const double *__restrict__ gn, const double *__restrict__ go,
                                                           long long n, double *__restrict__ s, double *__restrict__ y,
                                                           double *dots)
{
    __shared__ double red[LB_THREADS];
    double ys = 0.0, yy = 0.0, ss = 0.0;
    for (long long i = threadIdx.x; i < n; i += LB_THREADS) {
        const double sv = xn[i] - xo[i], yv = gn[i] - go[i];
        s[i] = sv;
        y[i] = yv;
        ys = fma(yv, sv, ys);
        yy = fma(yv, yv, yy);
        ss = fma(sv, sv, ss);
    }
    ys = lb_block_sum(ys, red);
    yy = lb_block_sum(yy, red);
    ss = lb_block_sum(ss, red);
    if (threadIdx.x == 0) {
        dots[0] = ys;
        dots[1] = yy;
        dots[2] = ss;
    }
}

// a . b (one workgroup, fixed summation order)
__global__ __launch_bounds__(LB_THREADS) void k_dot64(const double *__restrict__ a, const double *__restrict__ b,
                                                      long long n, double *out)
{
    __shared__ double red[LB_THREADS];
    double p = 0.0;
    for (long long i = threadIdx.x; i < n; i += LB_THREADS) p = fma(a[i], b[i], p);
    const double r = lb_block_sum(p, red);
    if (threadIdx.x == 0) *out = r;
}

// out = clip(x + t d, lo, hi) elementwise; lo / hi may be null (no bound)
__global__ void k_step_clip(const double *__restrict__ x, const double *__restrict__ d, double t,
                            const double *__restrict__ lo, const double *__restrict__ hi, long long n,
                            double *__restrict__ out)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        double v = fma(t, d[i], x[i]);
        if (lo && v < lo[i]) v = lo[i];
        if (hi && v > hi[i]) v = hi[i];
        out[i] = v;
    }
}

// v[i] = 0 on the binding set: entries at a bound whose gradient points out
// of the box (x <= lo with g > 0, or x >= hi with g < 0)
__global__ void k_bound_mask(const double *__restrict__ x, const double *__restrict__ g,
                             const double *__restrict__ lo, const double *__restrict__ hi, long long n,
                             double *__restrict__ v)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const bool at_lo = lo && x[i] <= lo[i] && g[i] > 0.0;
        const bool at_hi = hi && x[i] >= hi[i] && g[i] < 0.0;
        if (at_lo || at_hi) v[i] = 0.0;
    }
}

}  // namespace stfwi
