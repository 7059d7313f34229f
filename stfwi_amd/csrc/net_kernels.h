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

// Tiled form of the same convolution for cin (of this call) <= 64.  A block
// computes CONV_TROWS image rows x 64 columns of one image for a slice of up
// to 32 output channels; wave w owns row w of the tile (4 MFMA column tiles
// of 16 pixels).  Staged once per block in LDS, as BF16:
//   patch [CONV_TROWS + ks - 1][64 + ks - 1][ps]   the input rows with their
//         halo, channels innermost (cinp = cin rounded up to 8 of the ps
//         slots used), zero outside the image;
//   wp    [32][wps]   the weights in the GEMM's k order k = tap * cinp + ci,
//         Kp = ks*ks*cinp rounded up to 32 of the wps slots used, zero-padded;
//   boff  [Kp / 8]   patch offset of each group of 8 consecutive k.
// The row strides ps and wps are odd multiples of 4 banks (conv_strides), so
// the 16-byte operand reads of eight consecutive lanes hit disjoint banks.
// A B operand (8 consecutive k at one pixel) is then one 16-byte LDS read, an
// A operand one 16-byte read of wp.  Dynamic LDS; grid x = ceil(W / 64),
// y = ceil(H / CONV_TROWS), z = batch * ceil(cout / 32); 256 threads.
#define CONV_TROWS 4
__host__ __device__ inline void conv_strides(int cinp, int Kp, int &ps, int &wps)
{
    ps = cinp % 16 == 0 ? cinp + 8 : cinp;
    wps = Kp + 8;
}

// The weights in the tiled kernel's layout, BF16: wpk[ct][m][wps] holds
// output channel ct * 32 + m at k = tap * cinp + ci (zero where no weight).
// TR: the data-gradient weights (stored [cin][cout][ks][ks] of this call,
// taps flipped).  One thread per destination element.
template <int KS, bool TR>
__global__ void k_conv_pack(const float *__restrict__ wt, int cin, int cout, int cinp, int wps,
                            __bf16 *__restrict__ wpk, long long n)
{
    constexpr int K2 = KS * KS;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(e % wps);
        const long long row = e / wps;              // ct * 32 + m = output channel
        const int co = (int)row, tap = k / cinp, ci = k - tap * cinp;
        float v = 0.f;
        if (co < cout && tap < K2 && ci < cin)
            v = TR ? wt[((long long)ci * cout + co) * K2 + (K2 - 1 - tap)] : wt[((long long)co * cin + ci) * K2 + tap];
        wpk[e] = (__bf16)v;
    }
}

template <int KS>
__global__ __launch_bounds__(256) void k_conv_tile(const float *__restrict__ in, int cin, int H, int W,
                                                   const __bf16 *__restrict__ wpk, const float *__restrict__ bias,
                                                   int cout, float *__restrict__ out, int relu, int cinp, int Kp)
{
    constexpr int K2 = KS * KS, R = KS / 2, PR = CONV_TROWS + 2 * R, PC = 64 + 2 * R;
    extern __shared__ __attribute__((aligned(16))) unsigned char conv_smem[];
    int ps, wps;
    conv_strides(cinp, Kp, ps, wps);
    __bf16 *patch = reinterpret_cast<__bf16 *>(conv_smem);
    __bf16 *wp = patch + PR * PC * ps;
    int *boff = reinterpret_cast<int *>(wp + 32 * wps);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int x0 = blockIdx.x * 64, y0 = blockIdx.y * CONV_TROWS;
    const int nct = (cout + 31) / 32;
    const int b = blockIdx.z / nct, ct = blockIdx.z - b * nct;
    const int P = H * W;
    const float *inb = in + (long long)b * cin * P;
    // input patch: item = (channel group, pixel), pixels fastest; four
    // items' loads are in flight before their conversion and LDS store
    const int npix = PR * PC, ngrp = cinp / 8, nit = npix * ngrp;
    for (int base = tid; base < nit; base += 4 * 256) {
        float v[4][8];
        int dst[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int it = base + 256 * u;
            const int cg = it / npix, pi = it - cg * npix;
            const int pr = pi / PC, pc = pi - pr * PC;
            const int y = y0 - R + pr, x = x0 - R + pc;
            const bool ok = it < nit && y >= 0 && y < H && x >= 0 && x < W;
            dst[u] = it < nit ? pi * ps + cg * 8 : -1;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int ci = cg * 8 + j;
                v[u][j] = (ok && ci < cin) ? inb[(long long)ci * P + y * W + x] : 0.f;
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            nbf16x8 pk;
#pragma unroll
            for (int j = 0; j < 8; ++j) pk[j] = (__bf16)v[u][j];
            if (dst[u] >= 0) *reinterpret_cast<nbf16x8 *>(patch + dst[u]) = pk;
        }
    }
    // this slice's packed weights, 16 bytes per load
    {
        const uint4 *src = reinterpret_cast<const uint4 *>(wpk + (long long)ct * 32 * wps);
        uint4 *dst = reinterpret_cast<uint4 *>(wp);
        const int n16 = 32 * wps / 8;
        for (int i = tid; i < n16; i += 256) dst[i] = src[i];
    }
    // patch offset of every k group (groups past the last tap read tap 0:
    // their weights are zero)
    for (int g = tid; g < Kp / 8; g += 256) {
        int tap = (8 * g) / cinp;
        const int ci0 = 8 * g - tap * cinp;
        if (tap >= K2) tap = 0;
        const int ky = tap / KS, kx = tap - ky * KS;
        boff[g] = (ky * PC + kx) * ps + ci0;
    }
    __syncthreads();
    const int y = y0 + wave;
    if (y >= H) return;
    const int mt_n = min(2, (cout - ct * 32 + 15) / 16);
    const int grp = lane >> 4, col = lane & 15;
    const __bf16 *arow0 = wp + col * wps + 8 * grp;
    const __bf16 *arow1 = wp + (16 + col) * wps + 8 * grp;
    const __bf16 *bbase = patch + (wave * PC + col) * ps;
    // all four column tiles of the row per k step: one offset read, up to
    // two A and four B operand reads, then up to eight independent MFMAs
    nf32x4 acc[2][4];
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) acc[m][nt] = nf32x4{0.f, 0.f, 0.f, 0.f};
    for (int ks = 0; ks < Kp / 32; ++ks) {
        const int bo = boff[4 * ks + grp];
        nbf16x8 bb[4];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) bb[nt] = *reinterpret_cast<const nbf16x8 *>(bbase + nt * 16 * ps + bo);
        const nbf16x8 a0 = *reinterpret_cast<const nbf16x8 *>(arow0 + 32 * ks);
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) acc[0][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, bb[nt], acc[0][nt], 0, 0, 0);
        if (mt_n > 1) {
            const nbf16x8 a1 = *reinterpret_cast<const nbf16x8 *>(arow1 + 32 * ks);
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
                acc[1][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, bb[nt], acc[1][nt], 0, 0, 0);
        }
    }
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        const int x = x0 + nt * 16 + col;
        if (x >= W) continue;
        float *ob = out + (long long)b * cout * P + y * W + x;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int c0 = ct * 32 + 4 * grp + i, c1 = c0 + 16;
            if (c0 < cout) {
                float v = acc[0][nt][i] + (bias ? bias[c0] : 0.f);
                if (relu) v = v > 0.f ? v : 0.f;
                ob[(long long)c0 * P] = v;
            }
            if (mt_n > 1 && c1 < cout) {
                float v = acc[1][nt][i] + (bias ? bias[c1] : 0.f);
                if (relu) v = v > 0.f ? v : 0.f;
                ob[(long long)c1 * P] = v;
            }
        }
    }
}

// Weight-gradient partials over row chunks: chunk z covers rows
// [z * rg, (z + 1) * rg) of the batch's stacked rows (row index b * H + y);
// work[z][co][k] = sum over the chunk's pixels of dout[co][p] * col[k][p],
// one wave per 16 x 16 (co, k) tile, the pixels of a row taken 32 at a time
// in order.  Addressing is per row, so no division per element.
// grid: x = ceil(K / 16), y = ceil(cout / 16), z = ceil(batch * H / rg); 64 threads.
template <int KS>
__global__ __launch_bounds__(64) void k_conv_wgrad_rows(const float *__restrict__ in, int cin, int H, int W,
                                                        const float *__restrict__ dout, int cout, int batch, int rg,
                                                        float *__restrict__ work)
{
    constexpr int K2 = KS * KS, R = KS / 2;
    const int P = H * W;
    const int K = cin * K2;
    const int lane = threadIdx.x;
    const int k0 = blockIdx.x * 16, co0 = blockIdx.y * 16;
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
    const bool aok = am < cout, bok = bk < K;
    nf32x4 acc = {0.f, 0.f, 0.f, 0.f};
    const int row0 = blockIdx.z * rg, row1 = min(row0 + rg, batch * H);
    for (int row = row0; row < row1; ++row) {
        const int b = row / H, y = row - b * H;
        const float *arow = dout + ((long long)b * cout + (aok ? am : 0)) * P + (long long)y * W;
        const int yy = y + bky - R;
        const bool rok = bok && yy >= 0 && yy < H;
        const float *brow = in + ((long long)b * cin + bci) * P + (long long)(rok ? yy : 0) * W;
        for (int x0 = 0; x0 < W; x0 += 32) {
            float va[8], vb[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int x = x0 + kr + i, xx = x + bkx - R;
                va[i] = (aok && x < W) ? arow[x] : 0.f;
                vb[i] = (rok && x < W && xx >= 0 && xx < W) ? brow[xx] : 0.f;
            }
            nbf16x8 a, bb;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                a[i] = (__bf16)va[i];
                bb[i] = (__bf16)vb[i];
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bb, acc, 0, 0, 0);
        }
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

// bias-gradient partials: part[z][co] = sum of dout[b][co][p] over the
// flattened (batch, pixel) chunk z of `pch` elements; grid (cout, chunks),
// fixed-order tree within a block
__global__ __launch_bounds__(256) void k_bias_grad(const float *__restrict__ dout, int cout, int batch, int P, int pch,
                                                   float *__restrict__ part)
{
    __shared__ float red[256];
    const int co = blockIdx.x, z = blockIdx.y;
    const long long PT = (long long)batch * P, e0 = (long long)z * pch;
    float s = 0.f;
    for (long long e = e0 + threadIdx.x; e < e0 + pch && e < PT; e += 256) {
        const long long b = e / P, p = e - b * P;
        s += dout[(b * cout + co) * P + p];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(long long)z * cout + co] = red[0];
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
