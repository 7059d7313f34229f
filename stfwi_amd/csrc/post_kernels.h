// Gradient post-processing and preconditioner kernels (gfx950).
//   k_smooth_rows / k_smooth_cols  separable 'same'-mode zero-padded Gaussian
//                                  (fwi/loss.py:17-30; the normalised 2-D
//                                  kernel is the outer product of two
//                                  normalised 1-D kernels)
//   k_illum_hinv                   H = 1/(Ws + eps max Ws) / max(...)
//                                  (fwi/illumination.py:38-48)
//   k_eprecond3                    Plessix-Mulder weight
//                                  (fwi/preconditioning.py:107-139)
//   k_sos_traces                   DF-I biquad cascade + taper along t for
//                                  independent traces (fwi/filters.py)
#pragma once
#include <hip/hip_runtime.h>

namespace stfwi {

#define MAX_SMOOTH_R 64
struct SmoothW {
    double w[2 * MAX_SMOOTH_R + 1];
};

template <typename T>
__global__ void k_smooth_rows(const T *__restrict__ in, T *__restrict__ out, int nz, int nx, int r, SmoothW w)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x, z = blockIdx.y;
    if (c >= nx || z >= nz) return;
    double s = 0;
    for (int k = -r; k <= r; ++k) {
        const int cc = c + k;
        if (cc >= 0 && cc < nx) s += w.w[k + r] * (double)in[(long long)z * nx + cc];
    }
    out[(long long)z * nx + c] = (T)s;
}

template <typename T>
__global__ void k_smooth_cols(const T *__restrict__ in, T *__restrict__ out, int nz, int nx, int r, SmoothW w)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x, z = blockIdx.y;
    if (c >= nx || z >= nz) return;
    double s = 0;
    for (int k = -r; k <= r; ++k) {
        const int zz = z + k;
        if (zz >= 0 && zz < nz) s += w.w[k + r] * (double)in[(long long)zz * nx + c];
    }
    out[(long long)z * nx + c] = (T)s;
}

// sum over shots of the padded energy, crop [0:nz_c, pad:pad+nx_c] (the
// reference's crop assumes a free surface), then 1/(Ws + eps max)
// normalised.  One block; grids here are at most a few 1e5 cells.
template <typename T>
__global__ void k_illum_hinv(const T *__restrict__ ws, int n_shots, int nz_t, int nx_t, int z0, int pad, int nz_c,
                             int nx_c, double eps, T *out)
{
    __shared__ double red[256];
    const long long N = (long long)nz_c * nx_c, NT = (long long)nz_t * nx_t;
    double mx = 0;
    for (long long i = threadIdx.x; i < N; i += blockDim.x) {
        const long long r = i / nx_c, c = i % nx_c;
        double s = 0;
        for (int k = 0; k < n_shots; ++k) s += (double)ws[k * NT + (r + z0) * nx_t + c + pad];
        out[i] = (T)s;
        mx = fmax(mx, s);
    }
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    const double lam = eps * red[0];
    __syncthreads();
    double hmax = 0;
    for (long long i = threadIdx.x; i < N; i += blockDim.x) {
        const double h = 1.0 / ((double)out[i] + lam);
        out[i] = (T)h;
        hmax = fmax(hmax, h);
    }
    red[threadIdx.x] = hmax;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    const double hm = red[0];
    __syncthreads();
    for (long long i = threadIdx.x; i < N; i += blockDim.x) out[i] = (T)((double)out[i] / hm);
}

// eprecond3 on the padded grid of the summed energy, then crop to the
// physical window and normalise 1/We to max 1 (fwi/optim_lbfgsb.py:90-100).
// One block.
template <typename T>
__global__ void k_eprecond3(const T *__restrict__ ws, int n_shots, int nz_t, int nx_t, int z0, int pad, int nzd,
                            int nxd, int horizontal, double lmin, double lmax, double perp, double dx, double dz,
                            double eps, T *we_pad, T *out)
{
    __shared__ double red[256];
    const long long NT = (long long)nz_t * nx_t;
    double mx = 0;
    for (long long i = threadIdx.x; i < NT; i += blockDim.x) {
        const int r = (int)(i / nx_t), c = (int)(i % nx_t);
        double s = 0;
        for (int k = 0; k < n_shots; ++k) s += (double)ws[k * NT + i];
        const double cz = (double)(r - z0) * dz, cx = (double)(c - pad) * dx;
        const double cperp = horizontal ? cz : cx, cline = horizontal ? cx : cz, sp = horizontal ? dz : dx;
        const double y = fmax(fabs(cperp - perp), sp);
        double we = sqrt(s) * (asinh((lmax - cline) / y) - asinh((lmin - cline) / y));
        we = fabs(we);
        we_pad[i] = (T)we;
        mx = fmax(mx, we);
    }
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    const double add = eps * red[0];
    __syncthreads();
    // We = |sqrt(Ws) (...)| + eps max(We)
    for (long long i = threadIdx.x; i < NT; i += blockDim.x) we_pad[i] = (T)((double)we_pad[i] + add);
    __syncthreads();
    const long long N = (long long)nzd * nxd;
    double pm = 0;
    for (long long i = threadIdx.x; i < N; i += blockDim.x) {
        const long long r = i / nxd, c = i % nxd;
        const double p = 1.0 / (double)we_pad[(r + z0) * nx_t + c + pad];
        out[i] = (T)p;
        pm = fmax(pm, p);
    }
    red[threadIdx.x] = pm;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    const double m = red[0];
    __syncthreads();
    for (long long i = threadIdx.x; i < N; i += blockDim.x) out[i] = (T)((double)out[i] / m);
}

// independent traces x[t][trace]: taper the last n_taper samples then the
// DF-I cascade (forward), or the exact transpose (cascade in reverse time,
// sections in reverse order, then the taper)
template <typename T>
struct SosArgs {
    T *x;
    int nt, ntr, n_sos, n_taper, transpose;
    T coef[8][5];
};

template <typename T>
__global__ void k_sos_traces(SosArgs<T> a)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.ntr) return;
    const int n = a.n_taper < a.nt ? a.n_taper : a.nt;
    auto tw = [&](int t) -> T {
        if (n <= 0 || t < a.nt - n) return T(1);
        const int k = t - (a.nt - n) + 1;
        return (T)((cos((double)k / n * M_PI) + 1.0) / 2.0);
    };
    if (!a.transpose) {
        T x1[8], x2[8], y1[8], y2[8];
        for (int k = 0; k < a.n_sos; ++k) x1[k] = x2[k] = y1[k] = y2[k] = 0;
        for (int t = 0; t < a.nt; ++t) {
            T v = a.x[(long long)t * a.ntr + j];
            if (a.n_taper > 0) v = v * tw(t);
            for (int k = 0; k < a.n_sos; ++k) {
                const T *cf = a.coef[k];
                const T y = cf[0] * v + cf[1] * x1[k] + cf[2] * x2[k] - cf[3] * y1[k] - cf[4] * y2[k];
                x2[k] = x1[k]; x1[k] = v; y2[k] = y1[k]; y1[k] = y;
                v = y;
            }
            a.x[(long long)t * a.ntr + j] = v;
        }
    } else {
        T cx1[8], cx2[8], cy1[8], cy2[8];
        for (int k = 0; k < a.n_sos; ++k) cx1[k] = cx2[k] = cy1[k] = cy2[k] = 0;
        for (int t = a.nt - 1; t >= 0; --t) {
            T gcur = a.x[(long long)t * a.ntr + j];
            for (int k = a.n_sos - 1; k >= 0; --k) {
                const T *cf = a.coef[k];
                const T gy = gcur + cy1[k];
                const T gx = cx1[k] + cf[0] * gy;
                cx1[k] = cx2[k] + cf[1] * gy;
                cx2[k] = cf[2] * gy;
                cy1[k] = cy2[k] - cf[3] * gy;
                cy2[k] = -cf[4] * gy;
                gcur = gx;
            }
            if (a.n_taper > 0) gcur = gcur * tw(t);
            a.x[(long long)t * a.ntr + j] = gcur;
        }
    }
}

}  // namespace stfwi

namespace stfwi {

// Zhou imaging of given history cubes (nt x nz x nx each) on the padded grid:
// g_lambda = -dt sum_t (exx+ezz) div(B v_a), g_mu = -dt sum_t [2 exx d_x(B U_a)
// + 2 ezz d_z(B V_a) + exz (d_z(B U_a) + d_x(B V_a))], g_rho_lame =
// -dt sum_{t=1}^{nt-2} v(t) . (v_a(t+1) - v_a(t-1)) / 2dt, centred periodic
// differences (fwi/imaging.py:125-152).  One thread per cell walks t.
template <typename T>
__global__ void k_zhou_cubes(const T *__restrict__ ex, const T *__restrict__ ez, const T *__restrict__ es,
                             const T *__restrict__ ua, const T *__restrict__ va, const T *__restrict__ B,
                             const T *__restrict__ vx, const T *__restrict__ vz, int nt, int nz, int nx, T dx,
                             T dz, T dt, T *out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
    if (c >= nx || r >= nz) return;
    const long long N = (long long)nz * nx;
    const int cp = c + 1 == nx ? 0 : c + 1, cm = c == 0 ? nx - 1 : c - 1;
    const int rp = r + 1 == nz ? 0 : r + 1, rm = r == 0 ? nz - 1 : r - 1;
    const long long q = (long long)r * nx + c, qxp = (long long)r * nx + cp, qxm = (long long)r * nx + cm;
    const long long qzp = (long long)rp * nx + c, qzm = (long long)rm * nx + c;
    const T bxp = B[qxp], bxm = B[qxm], bzp = B[qzp], bzm = B[qzm];
    const T i2dx = T(2) * dx, i2dz = T(2) * dz, i2dt = T(2) * dt;
    T gl = 0, gm = 0, gr = 0;
    for (int t = 0; t < nt; ++t) {
        const long long o = (long long)t * N;
        const T dxBU = (bxp * ua[o + qxp] - bxm * ua[o + qxm]) / i2dx;
        const T dzBV = (bzp * va[o + qzp] - bzm * va[o + qzm]) / i2dz;
        const T dzBU = (bzp * ua[o + qzp] - bzm * ua[o + qzm]) / i2dz;
        const T dxBV = (bxp * va[o + qxp] - bxm * va[o + qxm]) / i2dx;
        gl += (ex[o + q] + ez[o + q]) * (dxBU + dzBV);
        gm += T(2) * ex[o + q] * dxBU + T(2) * ez[o + q] * dzBV + es[o + q] * (dzBU + dxBV);
        if (vx && t >= 1 && t <= nt - 2)
            gr += vx[o + q] * ((ua[o + N + q] - ua[o - N + q]) / i2dt) + vz[o + q] * ((va[o + N + q] - va[o - N + q]) / i2dt);
    }
    out[q] = -dt * gl;
    out[N + q] = -dt * gm;
    out[2 * N + q] = -dt * gr;
}

// ---------------------------------------------------------------------------
// eprecond1 (fwi/preconditioning.py:37-62): We = sqrt(Ws Wr) + eps max(...)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_sqrt_prod(const T *__restrict__ a, const T *__restrict__ b, long long n, T *__restrict__ out)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        out[i] = sqrt(a[i] * b[i]);
}

// part[blockIdx.x] = max (op = 1) or min (op = 0) of a grid-stride share of x
template <typename T>
__global__ __launch_bounds__(256) void k_minmax_part(const T *__restrict__ x, long long n, int op, T *__restrict__ part)
{
    __shared__ T red[256];
    T m = x[0];
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        m = op ? (x[i] > m ? x[i] : m) : (x[i] < m ? x[i] : m);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            const T o = red[threadIdx.x + h];
            red[threadIdx.x] = op ? (o > red[threadIdx.x] ? o : red[threadIdx.x]) : (o < red[threadIdx.x] ? o : red[threadIdx.x]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// x += (T)(eps * max), max = the single value at *mx
template <typename T>
__global__ void k_add_eps_max(T *__restrict__ x, long long n, const T *__restrict__ mx, double eps)
{
    const T add = (T)(eps * (double)*mx);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        x[i] = x[i] + add;
}

// ---------------------------------------------------------------------------
// taper_grad (fwi/preconditioning.py:144-183): depth taper, broadcast along x
// ---------------------------------------------------------------------------
__global__ void k_taper_grad(int nz, int nx, double dz, int gradt1, int gradt2, double exp_taper, int bottom, int gradb1,
                             int gradb2, double *__restrict__ out)
{
    const long long n = (long long)nz * nx;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const int j = (int)(e / nx);
        double t;
        if (j < gradt1) t = 0.0;
        else if (j <= gradt2) t = exp(-((double)(gradt2 - j) * (double)(gradt2 - j)));
        else t = pow(j * dz, exp_taper);
        if (bottom) {
            const double den = fmax((double)(gradb2 - gradb1), 1.0);
            if (j > gradb2) t = 0.0;
            else if (j >= gradb1) t *= exp(-((double)(j - gradb1) * (double)(j - gradb1)) / (den * den) * 4.0);
        }
        out[e] = t;
    }
}

// ---------------------------------------------------------------------------
// build_source_taper (fwi/taper.py:16-84)
// ---------------------------------------------------------------------------
// the normalised erf patch (ps x ps), one workgroup: erf(r / (sqrt 2 R)),
// minus its minimum, divided by the largest of its four edge-centre values,
// clipped to [0, 1]
__global__ __launch_bounds__(256) void k_src_patch(int ps, double dh, double srt_radius, double *__restrict__ patch)
{
    __shared__ double red[256];
    const double center = ps / 2.0, maxrad = sqrt(2.0) * srt_radius;
    double m = 1e300;
    for (int e = threadIdx.x; e < ps * ps; e += blockDim.x) {
        const int jj = e / ps, ii = e - jj * ps;
        const double x = (ii - center + 0.5) * dh, y = (jj - center + 0.5) * dh;
        const double v = erf(sqrt(x * x + y * y) / maxrad);
        patch[e] = v;
        m = fmin(m, v);
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + h]);
        __syncthreads();
    }
    const double pmin = red[0];
    __syncthreads();
    for (int e = threadIdx.x; e < ps * ps; e += blockDim.x) patch[e] -= pmin;
    __syncthreads();
    const int mid = ps / 2;
    const double ev = fmax(fmax(patch[mid], patch[mid * ps]), fmax(patch[mid * ps + ps - 1], patch[(ps - 1) * ps + mid]));
    __syncthreads();
    for (int e = threadIdx.x; e < ps * ps; e += blockDim.x) {
        const double v = patch[e] / ev;
        patch[e] = v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
    }
}

// taper[gz][gx] = min(1, min over sources whose patch covers the cell)
__global__ void k_src_stamp(int nz, int nx, int ps, const double *__restrict__ patch, const int *__restrict__ src,
                            int nsrc, double *__restrict__ out)
{
    const long long n = (long long)nz * nx;
    const int ijc = ps / 2;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const int gz = (int)(e / nx), gx = (int)(e - (long long)gz * nx);
        double t = 1.0;
        for (int s = 0; s < nsrc; ++s) {
            const int pi = gx - src[2 * s] + ijc, pj = gz - src[2 * s + 1] + ijc;
            if (pi >= 0 && pi < ps && pj >= 0 && pj < ps) t = fmin(t, patch[pj * ps + pi]);
        }
        out[e] = t;
    }
}

// taper = (taper - min) / max(taper - min) when that max is > 0, then zero
// the (2 filt + 1)^2 box around every source; mm = {min, max} of the input
__global__ void k_src_finish(int nz, int nx, const double *__restrict__ mm, const int *__restrict__ src, int nsrc,
                             int filt, double *__restrict__ out)
{
    const long long n = (long long)nz * nx;
    const double lo = mm[0], hi = mm[1] - mm[0];
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const int gz = (int)(e / nx), gx = (int)(e - (long long)gz * nx);
        double t = out[e] - lo;
        if (hi > 0) t = t / hi;
        for (int s = 0; s < nsrc; ++s)
            if (gz >= src[2 * s + 1] - filt && gz <= src[2 * s + 1] + filt && gx >= src[2 * s] - filt &&
                gx <= src[2 * s] + filt)
                t = 0.0;
        out[e] = t;
    }
}

}  // namespace stfwi
