// Kernels of the learned gradient preconditioner: a small U-Net whose
// convolutions run on the gfx950 matrix cores in BF16 with FP32
// accumulation (v_mfma_f32_16x16x32_bf16), plus the FP32 layer kernels
// around them.  No reference counterpart: the reference's plain method has
// no learned part.
//
// Activations are FP32 NCHW in memory; weights are FP32 masters
// [cout][cin][k][k].  Both are rounded to BF16 (round to nearest even) as
// they are loaded into MFMA operands.  A convolution is the implicit GEMM
//   out[co][p] = sum_k W[co][k] col[k][p],   k = (ci, ky, kx),  p = pixel,
// with M = cout, N = pixels, K = cin*ks*ks in chunks of 32.
//
// MFMA 16x16x32 bf16 lane layout (lane l of 64):
//   A (16 x 32):  row l % 16, k = 8 (l / 16) + i, i < 8
//   B (32 x 16):  col l % 16, k = 8 (l / 16) + i
//   D (16 x 16):  rows 4 (l / 16) + i (i < 4), col l % 16
#pragma once
#include <hip/hip_runtime.h>

namespace stfwi {

typedef __bf16 nbf16x8 __attribute__((ext_vector_type(8)));
typedef float nf32x4 __attribute__((ext_vector_type(4)));

// Convolution (ks = 1 or 3, stride 1, zero 'same' padding) of a batch.
// TR = true evaluates the data gradient: the weights are read transposed
// and flipped, W'[ci][co][ky][kx] = W[co][ci][ks-1-ky][ks-1-kx], where the
// stored tensor is [cin_of_this_call][cout_of_this_call][ks][ks].
// grid: x = ceil(P / 256) pixel blocks (4 waves x 64 pixels), y = ceil(cout / 16),
// z = batch; 256 threads.
template <int KS, bool TR>
__global__ __launch_bounds__(256) void k_conv_mfma(const float *__restrict__ in, int cin, int H, int W,
                                                   const float *__restrict__ wt, const float *__restrict__ bias,
                                                   int cout, float *__restrict__ out, int relu)
{
    constexpr int K2 = KS * KS, R = KS / 2;
    const int P = H * W;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int p0 = blockIdx.x * 256 + wave * 64;
    const int co0 = blockIdx.y * 16;
    const int b = blockIdx.z;
    const float *inb = in + (long long)b * cin * P;
    const int K = cin * K2;
    const int kr = 8 * (lane >> 4);
    const int am = co0 + (lane & 15);
    nf32x4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = nf32x4{0.f, 0.f, 0.f, 0.f};
    int py[4], px[4];
    bool pok[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int n = p0 + 16 * t + (lane & 15);
        pok[t] = n < P;
        py[t] = pok[t] ? n / W : 0;
        px[t] = pok[t] ? n - py[t] * W : 0;
    }
    for (int kc = 0; kc < K; kc += 32) {
        nbf16x8 a;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int k = kc + kr + i;
            float v = 0.f;
            if (am < cout && k < K) {
                const int ci = k / K2, r = k - ci * K2;
                if (!TR) {
                    v = wt[(long long)am * K + k];
                } else {
                    // stored [cin_call][cout_call] = [ci][am]: forward weights W[ci][am][..]
                    v = wt[((long long)ci * cout + am) * K2 + (K2 - 1 - r)];
                }
            }
            a[i] = (__bf16)v;
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            nbf16x8 bb;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int k = kc + kr + i;
                float v = 0.f;
                if (pok[t] && k < K) {
                    const int ci = k / K2, r = k - ci * K2;
                    const int ky = r / KS, kx = r - ky * KS;
                    const int yy = py[t] + ky - R, xx = px[t] + kx - R;
                    if (yy >= 0 && yy < H && xx >= 0 && xx < W) v = inb[(long long)ci * P + yy * W + xx];
                }
                bb[i] = (__bf16)v;
            }
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bb, acc[t], 0, 0, 0);
        }
    }
    float *ob = out + (long long)b * cout * P;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int n = p0 + 16 * t + (lane & 15);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int co = co0 + 4 * (lane >> 4) + i;
            if (co < cout && n < P) {
                float v = acc[t][i] + (bias ? bias[co] : 0.f);
                if (relu) v = v > 0.f ? v : 0.f;
                ob[(long long)co * P + n] = v;
            }
        }
    }
}

// Weight-gradient partials: work[z][co][k] = sum over the pixels of chunk z
// (batch x pixels flattened) of dout[co][p] * col[k][p].  One wave per
// 16 x 16 (co, k) tile and pixel chunk of `pch` pixels (a multiple of 32).
// grid: x = ceil(K / 16), y = ceil(cout / 16), z = chunks; 64 threads.
template <int KS>
__global__ __launch_bounds__(64) void k_conv_wgrad_mfma(const float *__restrict__ in, int cin, int H, int W,
                                                        const float *__restrict__ dout, int cout, int batch, int pch,
                                                        float *__restrict__ work)
{
    constexpr int K2 = KS * KS, R = KS / 2;
    const int P = H * W;
    const long long PT = (long long)batch * P;
    const int K = cin * K2;
    const int lane = threadIdx.x;
    const int k0 = blockIdx.x * 16, co0 = blockIdx.y * 16;
    const long long pbeg = (long long)blockIdx.z * pch;
    const int kr = 8 * (lane >> 4);
    const int am = co0 + (lane & 15);          // A row: output channel
    const int bk = k0 + (lane & 15);           // B column: (ci, ky, kx)
    int bci = 0, bky = 0, bkx = 0;
    if (bk < K) {
        bci = bk / K2;
        const int r = bk - bci * K2;
        bky = r / KS;
        bkx = r - bky * KS;
    }
    nf32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int pp = 0; pp < pch; pp += 32) {
        nbf16x8 a, bb;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const long long p = pbeg + pp + kr + i;
            float va = 0.f, vb = 0.f;
            if (p < PT) {
                const int b = (int)(p / P), n = (int)(p - (long long)b * P);
                if (am < cout) va = dout[((long long)b * cout + am) * P + n];
                if (bk < K) {
                    const int y = n / W, x = n - y * W;
                    const int yy = y + bky - R, xx = x + bkx - R;
                    if (yy >= 0 && yy < H && xx >= 0 && xx < W) vb = in[((long long)b * cin + bci) * P + yy * W + xx];
                }
            }
            a[i] = (__bf16)va;
            bb[i] = (__bf16)vb;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bb, acc, 0, 0, 0);
    }
    float *wz = work + (long long)blockIdx.z * cout * K;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int co = co0 + 4 * (lane >> 4) + i, k = k0 + (lane & 15);
        if (co < cout && k < K) wz[(long long)co * K + k] = acc[i];
    }
}

// out[i] = sum_z work[z][i] (fixed order), optionally added to out
__global__ void k_sum_chunks(const float *__restrict__ work, int nz, long long n, float *__restrict__ out, int add)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int z = 0; z < nz; ++z) s += work[(long long)z * n + i];
        out[i] = add ? out[i] + s : s;
    }
}

// db[co] = sum over batch and pixels of dout[b][co][p]; one workgroup per co
__global__ __launch_bounds__(256) void k_bias_grad(const float *__restrict__ dout, int cout, int batch, int P,
                                                   float *__restrict__ db)
{
    __shared__ float red[256];
    const int co = blockIdx.x;
    float s = 0.f;
    for (int b = 0; b < batch; ++b)
        for (int p = threadIdx.x; p < P; p += 256) s += dout[((long long)b * cout + co) * P + p];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[co] = red[0];
}

// g *= (y > 0): ReLU backward through its output
__global__ void k_relu_bwd(float *__restrict__ g, const float *__restrict__ y, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        if (!(y[i] > 0.f)) g[i] = 0.f;
}

// 2 x 2 average pooling of planes (C planes of H x W, H and W even)
__global__ void k_pool2(const float *__restrict__ in, long long planes, int H, int W, float *__restrict__ out)
{
    const int h = H / 2, w = W / 2;
    const long long n = planes * h * w;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const long long c = i / (h * w);
        const int r = (int)(i - c * h * w), y = r / w, x = r - y * w;
        const float *q = in + c * H * W + (2 * y) * W + 2 * x;
        out[i] = 0.25f * ((q[0] + q[1]) + (q[W] + q[W + 1]));
    }
}

__global__ void k_pool2_bwd(const float *__restrict__ gout, long long planes, int H, int W, float *__restrict__ gin)
{
    const int h = H / 2, w = W / 2;
    const long long n = planes * H * W;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const long long c = i / ((long long)H * W);
        const int r = (int)(i - c * H * W), y = r / W, x = r - y * W;
        gin[i] = 0.25f * gout[c * h * w + (y / 2) * w + x / 2];
    }
}

// nearest x2 upsampling of planes of h x w into planes of 2h x 2w at the
// given plane stride of the output (so it can write into a concat buffer)
__global__ void k_up2(const float *__restrict__ in, int planes, int h, int w, float *__restrict__ out,
                      long long out_plane_stride, int batch, long long in_batch_stride, long long out_batch_stride)
{
    const int H = 2 * h, W = 2 * w;
    const long long per = (long long)planes * H * W, n = per * batch;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int b = (int)(i / per);
        const long long j = i - b * per;
        const int c = (int)(j / ((long long)H * W));
        const int r = (int)(j - (long long)c * H * W), y = r / W, x = r - y * W;
        out[b * out_batch_stride + c * out_plane_stride + r] = in[b * in_batch_stride + (long long)c * h * w + (y / 2) * w + x / 2];
    }
}

__global__ void k_up2_bwd(const float *__restrict__ gout, int planes, int h, int w, float *__restrict__ gin,
                          long long gout_plane_stride, int batch, long long gin_batch_stride,
                          long long gout_batch_stride)
{
    const int W = 2 * w;
    const long long per = (long long)planes * h * w, n = per * batch;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int b = (int)(i / per);
        const long long j = i - b * per;
        const int c = (int)(j / ((long long)h * w));
        const int r = (int)(j - (long long)c * h * w), y = r / w, x = r - y * w;
        const float *q = gout + b * gout_batch_stride + c * gout_plane_stride + (2 * y) * W + 2 * x;
        gin[b * gin_batch_stride + j] = (q[0] + q[1]) + (q[W] + q[W + 1]);
    }
}

// Adam (Kingma & Ba 2015, Algorithm 1) on FP32 parameters
__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, long long n, float lr, float b1, float b2, float eps, float c1, float c2)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// masked mean squared error: g = 2 w (y - t) / sum(w); partial sums of
// w (y - t)^2 per workgroup into part[blockIdx.x]
__global__ __launch_bounds__(256) void k_mse(const float *__restrict__ y, const float *__restrict__ t,
                                             const float *__restrict__ wgt, long long n, float inv_wsum,
                                             float *__restrict__ g, float *__restrict__ part)
{
    __shared__ float red[256];
    float s = 0.f;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float wi = wgt ? wgt[i] : 1.f;
        const float d = y[i] - t[i];
        g[i] = 2.f * wi * d * inv_wsum;
        s += wi * d * d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

}  // namespace stfwi
