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

// Tiled form of the same convolution for one pass over up to 64 input
// channels (the host runs wider inputs as several passes).  A block
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
// output channel ct * 32 + m at k = tap * cinp + ci (zero where no weight),
// for the input channels [ci0, ci0 + cin) of the cin_all the weights hold.
// TR: the data-gradient weights (stored [cin_all][cout][ks][ks] of this
// call, taps flipped).  One thread per destination element.
template <int KS, bool TR>
__global__ void k_conv_pack(const float *__restrict__ wt, int ci0, int cin, int cin_all, int cout, int cinp, int wps,
                            __bf16 *__restrict__ wpk, long long n)
{
    constexpr int K2 = KS * KS;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(e % wps);
        const long long row = e / wps;              // ct * 32 + m = output channel
        const int co = (int)row, tap = k / cinp, ci = k - tap * cinp;
        float v = 0.f;
        if (co < cout && tap < K2 && ci < cin)
            v = TR ? wt[((long long)(ci0 + ci) * cout + co) * K2 + (K2 - 1 - tap)]
                   : wt[((long long)co * cin_all + ci0 + ci) * K2 + tap];
        wpk[e] = (__bf16)v;
    }
}

// A convolution with more input channels runs as passes over chunks of
// them: `in` points at the chunk's first channel, bstride is the input's
// elements per image, and accum = 1 adds the pass to the partial sums in
// `out` (the last pass alone gets the bias and the ReLU).
template <int KS>
__global__ __launch_bounds__(256) void k_conv_tile(const float *__restrict__ in, int cin, int H, int W,
                                                   const __bf16 *__restrict__ wpk, const float *__restrict__ bias,
                                                   int cout, float *__restrict__ out, int relu, int cinp, int Kp,
                                                   long long bstride, int accum)
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
    const float *inb = in + (long long)b * bstride;
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
                float v = acc[0][nt][i] + (accum ? ob[(long long)c0 * P] : 0.f) + (bias ? bias[c0] : 0.f);
                if (relu) v = v > 0.f ? v : 0.f;
                ob[(long long)c0 * P] = v;
            }
            if (mt_n > 1 && c1 < cout) {
                float v = acc[1][nt][i] + (accum ? ob[(long long)c1 * P] : 0.f) + (bias ? bias[c1] : 0.f);
                if (relu) v = v > 0.f ? v : 0.f;
                ob[(long long)c1 * P] = v;
            }
        }
    }
}

// Weight gradient: dw[co][k] = sum over pixels p of dout[co][p] col[k][p],
// k = (ci, ky, kx), as tiled BF16 MFMA GEMMs with FP32 accumulation.  The
// pixels are cut into tiles of WG_TR rows x 64 columns of one image; block
// (z, cg, mg) sums tiles z, z + gridDim.x, ... in that order into registers
// for the input channels [64 cg, 64 cg + 64) and output channels
// [32 mg, 32 mg + 32), then writes its partial into work[z][co][k] (the
// partials are summed in order by k_sum_chunks, so the result is
// deterministic).  Per tile the block stages dout as BF16 [co][pixel] and,
// per chunk of 16 input channels, the input rows with their halo as BF16
// [ci][row][position], tile column x at position x + 8.  An operand holds
// 8 consecutive pixels of one row: the aligned 16-byte read at the tap's
// unshifted position, and for kx - R = -1 / +1 the neighbouring dword,
// funnel-shifted by one element.  Wave w owns the N-tiles w, w + 4, ... of
// the group with both M-tiles, so each B operand feeds two MFMAs.  Every
// global load reads a clamped in-bounds address and selects zero outside
// the image, so the staging has no per-lane branches.  Row strides of 264
// and 88 halfwords are odd multiples of 16 bytes.
#define WG_TR 4
#define WG_AS 264
#define WG_BS 88
#define WG_NPART 512
template <int KS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_conv_wgrad_tile(const float *__restrict__ in, int cin, int H, int W,
                                                         const float *__restrict__ dout, int cout, int batch,
                                                         float *__restrict__ work)
{
    constexpr int K2 = KS * KS, R = KS / 2, BR = WG_TR + 2 * R;
    __shared__ __attribute__((aligned(16))) __bf16 As[32 * WG_AS];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[16 * BR * WG_BS];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int P = H * W, K = cin * K2;
    const int cg = blockIdx.y, mg = blockIdx.z;
    const int co0 = 32 * mg, nco = cout - co0 < 32 ? cout - co0 : 32;
    const int nmt = nco > 16 ? 2 : 1;
    const int ci0 = 64 * cg, nci = cin - ci0 < 64 ? cin - ci0 : 64;
    const int nchunk = (nci + 15) / 16;
    const int tiles_y = (H + WG_TR - 1) / WG_TR, tiles_x = (W + 63) / 64;
    const int ntiles = batch * tiles_y * tiles_x;
    // the lane's B column in each of the wave's N-tiles: LDS offset of tile
    // column 0 in staged row ky (-1 past the last column) and the tap's
    // column shift kx - R
    int boff[K2], bsh[K2];
    bool tvalid[K2];
#pragma unroll
    for (int u = 0; u < K2; ++u) {
        const int tl = wave + 4 * u;
        const int n = 16 * (cg * 4 * K2 + tl) + (lane & 15);
        tvalid[u] = tl / K2 < nchunk;
        boff[u] = -1;
        bsh[u] = 0;
        if (tvalid[u] && n < K) {
            const int ci = n / K2, rem = n - ci * K2;
            const int ky = rem / KS, kx = rem - ky * KS;
            boff[u] = ((ci & 15) * BR + ky) * WG_BS + 8;
            bsh[u] = kx - R;
        }
    }
    nf32x4 acc[2][K2];
#pragma unroll
    for (int u = 0; u < K2; ++u) acc[0][u] = acc[1][u] = nf32x4{0.f, 0.f, 0.f, 0.f};
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int b = tile / (tiles_y * tiles_x), r2 = tile - b * tiles_y * tiles_x;
        const int y0 = (r2 / tiles_x) * WG_TR, x0 = (r2 - (r2 / tiles_x) * tiles_x) * 64;
        __syncthreads();   // the previous tile's operand reads are done
        {
            // dout rows: thread (row j = wave, column lane), 16 channels at
            // a time, every load of a group issued before its stores
            const int y = y0 + wave, x = x0 + lane;
            const bool ok = y < H && x < W;
            const float *db = dout + ((long long)b * cout + co0) * P;
            const int po = ok ? y * W + x : 0;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                float va[16];
#pragma unroll
                for (int c = 0; c < 16; ++c) {
                    const int co = 16 * h + c;
                    const float t = db[(co < nco ? co : 0) * P + po];
                    va[c] = (ok && co < nco) ? t : 0.f;
                }
#pragma unroll
                for (int c = 0; c < 16; ++c) As[(16 * h + c) * WG_AS + wave * 64 + lane] = (__bf16)va[c];
            }
        }
        for (int cc = 0; cc < nchunk; ++cc) {
            if (cc > 0) __syncthreads();   // the previous chunk's B reads are done
            // rows y0 - R + row of channel ci0 + 16 cc + cil (pr = cil BR +
            // row): thread (wave, lane) takes the rows pr = wave, wave + 4,
            // ... at tile column lane - R, half of them at a time; thread t
            // then takes extra column e = t mod 2R (tile column 64 + e - R)
            // of row t / 2R
            constexpr int NPR = 4 * BR;
            const float *cb = in + ((long long)b * cin + ci0 + 16 * cc) * P;
            const int nrow = nci - 16 * cc;   // channels present in this chunk
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                float vb[NPR / 2];
#pragma unroll
                for (int m = 0; m < NPR / 2; ++m) {
                    const int pr = wave + 4 * (h * (NPR / 2) + m), cil = pr / BR, row = pr - cil * BR;
                    const int y = y0 + row - R, xa = x0 - R + lane;
                    const bool oka = cil < nrow && y >= 0 && y < H && xa >= 0 && xa < W;
                    const float t = cb[oka ? cil * P + y * W + xa : 0];
                    vb[m] = oka ? t : 0.f;
                }
#pragma unroll
                for (int m = 0; m < NPR / 2; ++m) {
                    const int pr = wave + 4 * (h * (NPR / 2) + m);
                    Bs[pr * WG_BS + lane - R + 8] = (__bf16)vb[m];
                }
            }
            if (R > 0) {
                const bool act = tid < 16 * BR * 2 * R;
                const int pr = act ? tid / (2 * R) : 0, e = tid - (tid / (2 * R)) * (2 * R);
                const int cil = pr / BR, row = pr - cil * BR;
                const int y = y0 + row - R, xe = x0 - R + 64 + e;
                const bool oke = act && cil < nrow && y >= 0 && y < H && xe < W;
                const float t = cb[oke ? cil * P + y * W + xe : 0];
                Bs[pr * WG_BS + (act ? 64 + e - R + 8 : 80 + (lane & 7))] = (__bf16)(oke ? t : 0.f);
            }
            __syncthreads();
#pragma unroll 1
            for (int ks = 0; ks < 2 * WG_TR; ++ks) {
                const int j = ks >> 1, xb = (ks & 1) * 32 + 8 * (lane >> 4);
                const nbf16x8 a0 = *reinterpret_cast<const nbf16x8 *>(As + (lane & 15) * WG_AS + j * 64 + xb);
                const nbf16x8 a1 = *reinterpret_cast<const nbf16x8 *>(As + (16 + (lane & 15)) * WG_AS + j * 64 + xb);
#pragma unroll
                for (int u = 0; u < K2; ++u) {
                    if (!tvalid[u] || (wave + 4 * u) / K2 != cc) continue;
                    const int base = (boff[u] >= 0 ? boff[u] : 8) + j * WG_BS + xb;
                    uint4 q = *reinterpret_cast<const uint4 *>(Bs + base);
                    if (R > 0) {
                        const unsigned pv = *reinterpret_cast<const unsigned *>(Bs + base - 2);
                        const unsigned nv = *reinterpret_cast<const unsigned *>(Bs + base + 8);
                        const unsigned s0 = __builtin_amdgcn_alignbyte(q.x, pv, 2);
                        const unsigned s1 = __builtin_amdgcn_alignbyte(q.y, q.x, 2);
                        const unsigned s2 = __builtin_amdgcn_alignbyte(q.z, q.y, 2);
                        const unsigned s3 = __builtin_amdgcn_alignbyte(q.w, q.z, 2);
                        const unsigned s4 = __builtin_amdgcn_alignbyte(nv, q.w, 2);
                        if (bsh[u] < 0) q = uint4{s0, s1, s2, s3};
                        else if (bsh[u] > 0) q = uint4{s1, s2, s3, s4};
                    }
                    if (boff[u] < 0) q = uint4{0u, 0u, 0u, 0u};
                    const nbf16x8 bb = __builtin_bit_cast(nbf16x8, q);
                    acc[0][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, bb, acc[0][u], 0, 0, 0);
                    if (nmt > 1) acc[1][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, bb, acc[1][u], 0, 0, 0);
                }
            }
        }
    }
    float *wz = work + (long long)blockIdx.x * cout * K;
#pragma unroll
    for (int u = 0; u < K2; ++u) {
        const int k = 16 * (cg * 4 * K2 + wave + 4 * u) + (lane & 15);
        if (!tvalid[u] || k >= K) continue;
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int co = 16 * mt + 4 * (lane >> 4) + i;
                if (mt < nmt && co < nco) wz[(long long)(co0 + co) * K + k] = acc[mt][u][i];
            }
    }
}

// out[i] = sum_z work[z][i], optionally added to out.  256 threads per 64
// outputs: quarter q of the threads sums z in [q nz / 4, (q + 1) nz / 4) in
// order, then the four quarter sums are added in order (a fixed order, so
// the result is deterministic).  grid: ceil(n / 64) blocks of 256 threads.
__global__ __launch_bounds__(256) void k_sum_chunks(const float *__restrict__ work, int nz, long long n,
                                                    float *__restrict__ out, int add)
{
    __shared__ float red[4][64];
    const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
    const long long i = (long long)blockIdx.x * 64 + lane;
    const int z0 = (int)((long long)nz * q / 4), z1 = (int)((long long)nz * (q + 1) / 4);
    float sum = 0.f;
    if (i < n) {
#pragma unroll 8
        for (int z = z0; z < z1; ++z) sum += work[(long long)z * n + i];
    }
    red[q][lane] = sum;
    __syncthreads();
    if (q == 0 && i < n) {
        const float t = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
        out[i] = add ? out[i] + t : t;
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
