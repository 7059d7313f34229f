// Device kernels of the elastic velocity-strain FWI gradient path (gfx950).
//
// Layout (all fields row-major, z = row):
//   * every wavefield/modulus buffer is (nz + 8) x ld with ld >= nx + 8 and a
//     4-row zero margin above and below; the row gap [nx, ld) is never
//     written, so stencil loads at +-4 cells stay inside the allocation and
//     read zeros there.  `origin` = 4*ld is the offset of cell (0,0).
//   * C-PML memories exist only in the PML strips: x-strips are nz x 2*pad
//     (left pad columns, then right pad columns), z-strips are zrows rows of
//     ld elements (the field pitch, so that a tile's 64 columns are two whole
//     128-byte lines; bottom pad rows first, then the top pad rows when
//     there is no free surface).  Outside the strips the reference's recursion is an exact
//     no-op (b = 1, a = 0, kappa = 1, psi stays 0).
//   * histories keep the physical cells only: the strain history as
//     [frame][field][nz_dom][hpitch] with physical column cp at cp + hcoff
//     (hcoff = pad mod 32, so that the 64 columns of a tile start on a
//     128-byte line), the velocity history as [frame][field][nz_dom*nx_dom].
//
// Every derivative keeps the reference's per-term output ranges
// (fwi/forward.py:256-441): term 1 covers its own slice, term k >= 2 the
// range where both reads are in bounds ([k-1, n-1-k] for D+, [k, n-k] for
// D-).  The adjoint kernels are the exact transposes written as gathers:
// every output cell sums the contributions that the reference scatters into
// it (fwi/adjoint_jax_kernel.py:208-455), so no atomics are needed.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace stfwi {

typedef long long i64;

template <int FD> struct FDC;
template <> struct FDC<2> { static constexpr int NK = 1; };
template <> struct FDC<4> { static constexpr int NK = 2; };
template <> struct FDC<8> { static constexpr int NK = 4; };

// staggered-grid coefficients (fwi/forward.py:177-182)
template <int FD>
__host__ __device__ constexpr double fdc(int k)
{
    return FD == 8 ? (k == 1 ? 1225.0 / 1024.0 : k == 2 ? -245.0 / 3072.0 : k == 3 ? 49.0 / 5120.0 : -5.0 / 7168.0)
         : FD == 4 ? (k == 1 ? 9.0 / 8.0 : -1.0 / 24.0)
                   : 1.0;
}

struct ShotDev {
    int sx, sz;        // padded-grid source cell
    int type;          // 0 force, 1 moment
    int nt;            // record length
    int hoff;          // first history frame of the shot (frames packed by shot)
    int rec_off, nrec; // receivers [rec_off, rec_off + nrec) of rec_cell
    int uc_off, nuc;   // merged receiver cells [uc_off, uc_off + nuc)
};

template <typename T>
struct Geo {
    int nz, nx, ld;          // padded grid and field pitch
    int z0, pad, nzd, nxd;   // physical window
    int fs, zrows;           // free surface; rows of a z-strip
    i64 fstride;             // elements per field buffer
    i64 origin;              // offset of cell (0,0) in a field buffer
    i64 np;                  // physical cells
    int hpitch, hcoff;       // strain history rows: physical column cp at cp + hcoff of hpitch
    i64 hnp;                 // strain history elements per field (nzd * hpitch)
    T dt, dx, dz, dtdx, dtdz, hdtdx, hdtdz, dzdx, fsrc;
};

template <typename T>
__device__ __forceinline__ int xstrip(const Geo<T> &g, int r, int c)
{
    if (c < g.pad) return r * 2 * g.pad + c;
    if (c >= g.nx - g.pad) return r * 2 * g.pad + c - g.nx + 2 * g.pad;
    return -1;
}

template <typename T>
__device__ __forceinline__ int zstrip(const Geo<T> &g, int r, int c)
{
    if (r >= g.nz - g.pad) return (r - (g.nz - g.pad)) * g.ld + c;
    if (!g.fs && r < g.pad) return (g.pad + r) * g.ld + c;
    return -1;
}

// Global accesses of the time-step kernels.  ldg/stg: raw buffer loads and
// stores through a resource built from a block-uniform base pointer (kept in
// SGPRs) with a 32-bit byte offset per lane, so no per-lane 64-bit address
// arithmetic is emitted.  The resource's range is unbounded: every index
// passed is >= 0 and inside the allocation.  Word 3 is the CDNA (gfx9
// family) raw-buffer configuration.
//
// The wavefield-state stores (stg, stgg) carry the non-temporal hint: the
// next kernel reads the state after the kernel boundary, whose acquire
// invalidates L2, so a state line kept dirty in L2 only adds to the
// write-back at the end of the kernel.  The exact-imaging gradient
// accumulators (rows in the history's line-aligned layout) and the z-strip
// C-PML memories (rows of the field pitch) are stored the same way; the
// x-strip memories keep plain stores (a tile row holds 20 of their cells,
// less than a line, and with the hint they were slower).
#define STFWI_BUF_WORD3 0x00020000
// cache-policy operand of the buffer stores: non-temporal
#define STFWI_BUF_NT 2

template <typename T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const T *base)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<T *>(base), (short)0, -1, STFWI_BUF_WORD3);
}

__device__ __forceinline__ float buf_ld(__amdgpu_buffer_rsrc_t r, const float *, int idx)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, idx * 4, 0, 0));
}

__device__ __forceinline__ double buf_ld(__amdgpu_buffer_rsrc_t r, const double *, int idx)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, idx * 8, 0, 0));
}

__device__ __forceinline__ void buf_st(__amdgpu_buffer_rsrc_t r, float v, int idx)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, idx * 4, 0, STFWI_BUF_NT);
}

__device__ __forceinline__ void buf_st(__amdgpu_buffer_rsrc_t r, double v, int idx)
{
    typedef unsigned int u2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, idx * 8, 0, STFWI_BUF_NT);
}

template <typename T>
__device__ __forceinline__ T ldg(const T *base, int idx)
{
    return buf_ld(buf_rsrc(base), base, idx);
}

template <typename T>
__device__ __forceinline__ void stg(T *base, int idx, T v)
{
    buf_st(buf_rsrc(base), v, idx);
}

// ldgg/stgg: the same accesses as plain global loads and stores with a
// 32-bit unsigned byte offset from the uniform base.  The velocity-forward
// and strain-adjoint kernels use the buffer form, the strain-forward and
// velocity-adjoint kernels this one (the faster of the two for each, as
// timed on an MI355X).  The offset passes through an empty asm statement
// (no instruction; it only hides how the offset was computed), so that the
// compiler keeps it as a 32-bit VGPR and selects the SGPR-base addressing
// form instead of a 64-bit vector add per access.
__device__ __forceinline__ uint32_t opaque_off(uint32_t q)
{
    asm("" : "+v"(q));
    return q;
}

template <typename T>
__device__ __forceinline__ T ldgg(const T *base, int idx)
{
    const uint32_t q = opaque_off((uint32_t)idx * (uint32_t)sizeof(T));
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + q);
}

template <typename T>
__device__ __forceinline__ void stgg(T *base, int idx, T v)
{
    const uint32_t q = opaque_off((uint32_t)idx * (uint32_t)sizeof(T));
    __builtin_nontemporal_store(v, reinterpret_cast<T *>(reinterpret_cast<char *>(base) + q));
}

// stst: store of a write-once stream (the forward history, snapshots) with
// the non-temporal hint, so that these streams do not displace the resident
// wavefield state from the caches; they are read back many steps later.
// base is block-uniform and idx a 32-bit offset from it, as for stgg.
template <typename T>
__device__ __forceinline__ void stst(T *base, int idx, T v)
{
    const uint32_t q = opaque_off((uint32_t)idx * (uint32_t)sizeof(T));
    __builtin_nontemporal_store(v, reinterpret_cast<T *>(reinterpret_cast<char *>(base) + q));
}

// profile layout per axis: kappa[n], b[n], a[n], q[n], r[n] with
// q = 1/kappa + a and r = 1/kappa.  The forward recursion multiplies by r
// where the reference divides by kappa (one rounding apart; no divide
// sequence in the edge tiles).
template <typename T>
__device__ __forceinline__ T pmlT_val(T G, T psi_old, const T *prof, int j, int n)
{
    return fma(prof[3 * n + j], G, prof[2 * n + j] * psi_old); // (1/kappa + a) G + a psi
}

template <typename T>
__device__ __forceinline__ T pmlT_psi(T G, T psi_old, const T *prof, int j, int n)
{
    return prof[n + j] * (G + psi_old);                        // b (G + psi)
}

// pmlT_val with the profile values and the memory already loaded
template <typename T>
__device__ __forceinline__ T pmlT_pre(T G, T psi_old, T q, T a)
{
    return fma(q, G, a * psi_old);
}

// C-PML inputs of the transposed recursion at one staged cell: the two old
// memories and the (1/kappa + a, a) profile pair.  Cells outside the strip
// (i < 0) read entry 0, which a select discards afterwards.
template <typename T>
struct PmlIn {
    T m0, m1, q, a, b;
    // BUF selects buffer (ldg) or global (ldgg) accesses
    template <bool BUF>
    __device__ __forceinline__ void load(const T *mem0, const T *mem1, int i, const T *prof, int j, int n)
    {
        const int ic = i >= 0 ? i : 0;
        const int jc = j < 0 ? 0 : (j >= n ? n - 1 : j);
        auto ld = [](const T *b_, int k) { return BUF ? ldg(b_, k) : ldgg(b_, k); };
        m0 = ld(mem0, ic);
        m1 = ld(mem1, ic);
        q = ld(prof + 3 * n, jc);
        a = ld(prof + 2 * n, jc);
        b = ld(prof + n, jc);
    }
    // the two memories only; the profile values copied from `k`, loaded
    // once for all the cells of one column
    template <bool BUF>
    __device__ __forceinline__ void load_mem(const T *mem0, const T *mem1, int i, const PmlIn &k)
    {
        const int ic = i >= 0 ? i : 0;
        auto ld = [](const T *b_, int kk) { return BUF ? ldg(b_, kk) : ldgg(b_, kk); };
        m0 = ld(mem0, ic);
        m1 = ld(mem1, ic);
        q = k.q;
        a = k.a;
        b = k.b;
    }
    // the profile values of column (or row) j only
    template <bool BUF>
    __device__ __forceinline__ void load_coef(const T *prof, int j, int n)
    {
        const int jc = j < 0 ? 0 : (j >= n ? n - 1 : j);
        auto ld = [](const T *b_, int kk) { return BUF ? ldg(b_, kk) : ldgg(b_, kk); };
        q = ld(prof + 3 * n, jc);
        a = ld(prof + 2 * n, jc);
        b = ld(prof + n, jc);
    }
};

// the forward recursion's profile values (b, a, 1/kappa) at one column or
// row, loaded once for all the derivatives that share it
template <typename T>
struct PmlFwd {
    T b, a, r;
};

template <typename T>
__device__ __forceinline__ PmlFwd<T> pml_fwd_coef(const T *prof, int j, int n)
{
    return PmlFwd<T>{prof[n + j], prof[2 * n + j], prof[4 * n + j]};
}

// the forward recursion psi' = b psi + a D, D' = D / kappa + psi' (the
// division as a multiplication by 1/kappa)
template <typename T>
__device__ __forceinline__ T pml_fwd_apply(T D, T psi_old, T *psi, const PmlFwd<T> &k)
{
    const T pn = fma(k.b, psi_old, k.a * D);
    *psi = pn;
    return D * k.r + pn;
}

// the same for a z-strip memory, stored with the non-temporal hint
template <typename T>
__device__ __forceinline__ T pml_fwd_apply_nt(T D, T psi_old, T *psi, const PmlFwd<T> &k)
{
    const T pn = fma(k.b, psi_old, k.a * D);
    __builtin_nontemporal_store(pn, psi);
    return D * k.r + pn;
}

// ---------------------------------------------------------------------------
// model set-up (fwi/forward.py:193-238)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_model(Geo<T> g, const T *__restrict__ vs, const T *__restrict__ vp,
                        const T *__restrict__ rho, const T *__restrict__ bimg_phys,
                        T *L2M, T *M, T *B, T *Bimg, T *Lr0)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (c >= g.nx || r >= g.nz) return;
    // nearest-neighbour extension into the PML == clamped indexing
    int rr = r - g.z0, cc = c - g.pad;
    rr = rr < 0 ? 0 : (rr >= g.nzd ? g.nzd - 1 : rr);
    cc = cc < 0 ? 0 : (cc >= g.nxd ? g.nxd - 1 : cc);
    const i64 ip = (i64)rr * g.nxd + cc;
    const T s = vs[ip], p = vp[ip], d = rho[ip];
    const T m = d * (s * s), l2m = d * (p * p);
    const i64 q = g.origin + (i64)r * g.ld + c;
    M[q] = m;
    L2M[q] = l2m;
    B[q] = T(1) / d;
    Bimg[q] = bimg_phys ? bimg_phys[ip] : T(1) / d;
    if (r == 0) Lr0[c] = -(l2m - T(2) * m) / l2m;   // L_ratio = -lambda / (lambda + 2 mu)
}

template <typename T>
__global__ void k_vpmax(const T *__restrict__ vp, i64 n, double *out)
{
    __shared__ double red[256];
    double m = -1e300;
    for (i64 i = threadIdx.x; i < n; i += blockDim.x) m = fmax(m, (double)vp[i]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// C-PML profiles (fwi/forward.py:35-104): N = 2, Rc = 1e-3, L = pad*dx for
// both axes, d = d0 dist^2, kappa = 1 + dist^2, alpha = pi fc (1 - dist).
template <typename T>
__global__ void k_profiles(int nz, int nx, int pad, int fs, int nzd, double dx, double dt, double fc,
                           const double *vpmax, T *profx, T *profz)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const double Lp = pad * dx;
    const double d0 = -(2 + 1) * (*vpmax) * log(1e-3) / (2 * Lp);
    for (int axis = 0; axis < 2; ++axis) {
        const int n = axis == 0 ? nx : nz;
        if (j >= n) continue;
        int i = -1;   // profile index of this column / row
        if (axis == 0) {
            if (j < pad) i = pad - 1 - j;
            else if (j >= nx - pad) i = j - (nx - pad);
        } else {
            const int bottom = fs ? nzd : pad + nzd;
            if (j >= bottom && j < bottom + pad) i = j - bottom;
            else if (!fs && j < pad) i = pad - 1 - j;
        }
        double d = 0, k = 1, al = 0;
        if (i >= 0) {
            const double dist = (i + 1) * dx / Lp;
            d = d0 * (dist * dist);
            k = 1.0 + (2.0 - 1.0) * (dist * dist);
            al = M_PI * fc * (1.0 - dist);
        }
        const double b = exp(-(d / k + al) * dt);
        const double den = k * (d + k * al);
        const double a = den > 1e-10 ? d * (b - 1.0) / den : 0.0;
        T *p = axis == 0 ? profx : profz;
        p[j] = (T)k;
        p[n + j] = (T)b;
        p[2 * n + j] = (T)a;
        p[3 * n + j] = T(1) / (T)k + (T)a;
        p[4 * n + j] = T(1) / (T)k;
    }
}

template <typename T>
__global__ void k_fill(T *p, i64 n, T v)
{
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) p[i] = v;
}

// ---------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------
// Division by a launch-constant divisor without a divide sequence
// (Granlund and Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994, Fig. 4.1): for 1 <= d < 2^31 and
// l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1; then for every 32-bit
// n, n / d = (t + ((n - t) >> s1)) >> s2 with t = mulhi(m, n), s1 = min(l, 1)
// and s2 = max(l - 1, 0).  The host builds it once per launch; in a kernel
// it is five scalar instructions.
struct UDiv {
    unsigned m, s1, s2;
};

inline UDiv make_udiv(unsigned d)
{
    unsigned l = 0;
    while ((1ull << l) < d) ++l;
    UDiv u;
    u.m = (unsigned)((((1ull << 32) * ((1ull << l) - d)) / d) + 1);
    u.s1 = l < 1 ? l : 1;
    u.s2 = l > 1 ? l - 1 : 0;
    return u;
}

__device__ __forceinline__ unsigned udiv(unsigned n, const UDiv &u)
{
    const unsigned t = __umulhi(u.m, n);
    return (t + ((n - t) >> u.s1)) >> u.s2;
}

template <typename T>
struct FwdArgs {
    Geo<T> g;
    const ShotDev *shots;
    const T *L2M, *M, *B, *Lr0;      // moduli buffers (origin applied by the kernel)
    T *state;                        // per shot: U V exx ezz exz
    i64 sstride;
    T *psix, *psiz;                  // per shot: 5 x-strip / 5 z-strip arrays
    i64 psix_stride, psiz_stride, psix_n, psiz_n;
    const T *profx, *profz;
    const T *wav;                    // [shot][nt_max]
    int nt_max;
    T *rec;                          // [shot][4][nt_max][nrec_max]
    int nrec_max, rec_mask;
    const i64 *rec_cell;             // cell offset r*ld + c of every receiver
    T *hist;                         // [frame][3][hnp] strains (rows of hpitch), shot s from frame hoff
    T *hist2;                        // [frame][2][np]: 1 = dU/dB,dV/dB  2 = U,V
    int hist_nt, hist_t0;            // frames per shot with checkpoints; step of frame 0
    int hist2_mode;
    T *dxu0;                         // [shot][nt_max][nx]
    T *illum;                        // [shot][nz*nx]
    int illum_comp;
    T *snap;                         // [shot][nsnap][nt_max][nz*nx]
    int snap_mask, nsnap;
    int tiles_x, tiles_z, ntiles;
    int shot0;                         // first shot of this launch's group
    int nt_min;                        // shortest record among the launch's shots
    UDiv dgx, dtx;                     // division by gridDim.x and by tiles_x
    int lean;                          // the LEAN instantiations may be used
};

#define TILE_X 64
#define BLOCK_Y 4

// Each time-step kernel also has a LEAN instantiation for the common case
// of a gradient (or a forward) without the optional outputs: the launch
// picks it when the arguments already hold the values it fixes, so both
// instantiations compute the same thing; fixing them at compile time drops
// the unused paths and frees the scalar registers they held.
//
// Every time-step kernel works on a TZ x TILE_X tile (TZ = 8, 16 or 32, a
// template parameter the autotuner picks) with 256 threads (64 x 4, each
// thread owns TZ/4 rows).  Phase 1 stages what the stencils read (stress
// products, velocities, or the C-PML-transposed derivative adjoints) for
// the tile plus a 4-cell halo in LDS, computing each staged value once: a
// z-slab of TZ+8 rows over the tile's columns (whose middle rows also fill
// the tile part of the x-slab) and the 4+4 halo columns of the x-slab.
// Phase 2 evaluates the per-term-masked stencils from LDS.  A tile whose
// columns lie 8 cells inside the x-PML strips is "x-interior" and one whose
// rows lie 8 cells inside the z-PML strips and below the free-surface row is
// "z-interior"; the masks and C-PML work along an interior axis are compiled
// out (four instantiations of each body).

// XCD-aware block order: workgroup n is dispatched to XCD n % 8, so give
// each XCD a contiguous range of logical blocks; neighbouring tiles (which
// share halos) and whole shots then stay in one XCD's L2.  The map is a
// bijection on [0, gridDim.x * gridDim.z) whatever the dispatch order.
// dgx divides by gridDim.x.
__device__ __forceinline__ void logical_block(const UDiv &dgx, int &bx, int &bz)
{
    const int gx = gridDim.x;
    const int n = gx * gridDim.z;
    const int L = blockIdx.x + gx * blockIdx.z;
    const int q = n >> 3, rmd = n & 7;
    const int xcd = L & 7, w = L >> 3;
    // XCDs below rmd hold q + 1 logical blocks, the others q (no branch)
    const int lg = xcd * q + min(xcd, rmd) + w;
    bz = (int)udiv((unsigned)lg, dgx);
    bx = lg - bz * gx;
}

template <typename T>
__device__ __forceinline__ bool interior_x(const Geo<T> &g, int c0)
{
    return c0 >= g.pad + 8 && c0 + TILE_X + 8 <= g.nx - g.pad;
}

template <typename T>
__device__ __forceinline__ bool interior_z(const Geo<T> &g, int r0, int tz)
{
    const int top = g.fs ? 8 : g.pad + 8;
    return r0 >= top && r0 + tz + 8 <= g.nz - g.pad;
}

// The halo cells of a tile are staged by the block's threads in rounds:
// round k gives entry e = tid + 256 k of the nh halo cells to thread tid.  A
// wavefront (the 64 threads of one threadIdx.y) whose first entry of a round
// is past the halo has no valid lane in it and skips the round's loads,
// C-PML inputs and LDS stores with a scalar branch: threadIdx.y is uniform
// across the wavefront.  Lanes past the halo in a wavefront that does have
// valid lanes keep their selects (they load cell 0 and store into a spare
// LDS row).
__device__ __forceinline__ bool halo_round(int k, int nh)
{
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.y);
    return 256 * k + TILE_X * w < nh;
}

#define CX(x) (IX || (x))
#define CZ(x) (IZ || (x))
#define STFWI_DISPATCH(ix, iz, body)                                          \
    do {                                                                      \
        if (ix) {                                                             \
            if (iz) body(std::true_type{}, std::true_type{});                 \
            else body(std::true_type{}, std::false_type{});                   \
        } else {                                                              \
            if (iz) body(std::false_type{}, std::true_type{});                \
            else body(std::false_type{}, std::false_type{});                  \
        }                                                                     \
    } while (0)

// velocity half-step (fwi/forward.py:253-367) + exx/ezz receivers of the
// previous step (strains are read-only here)
template <typename T, int FD, int TZ, bool LEAN>
__global__ __launch_bounds__(256) void k_fwd_vel(FwdArgs<T> a, int it)
{
    if constexpr (LEAN) {
        // no velocity history
        a.hist2_mode = 0;
    }
    const Geo<T> &g = a.g;
    int bx, s;
    logical_block(a.dgx, bx, s);
    s += a.shot0;
    const ShotDev sh = a.shots[s];
    if (bx >= a.ntiles) {
        const int j = (bx - a.ntiles) * 256 + threadIdx.y * TILE_X + threadIdx.x;
        const int itp = it - 1;
        if (itp < 0 || itp >= sh.nt || j >= sh.nrec || !a.rec) return;
        const i64 cell = g.origin + a.rec_cell[sh.rec_off + j];
        const T *st = a.state + s * a.sstride;
        T *ro = a.rec + (i64)s * 4 * a.nt_max * a.nrec_max + (i64)itp * a.nrec_max + j;
        if (a.rec_mask & 4) ro[(i64)2 * a.nt_max * a.nrec_max] = st[2 * g.fstride + cell];
        if (a.rec_mask & 8) ro[(i64)3 * a.nt_max * a.nrec_max] = st[3 * g.fstride + cell];
        return;
    }
    // a shot past its record has no work at this step.  Only launches at or
    // past the shortest record among their shots test it: the others
    // issue their staging loads without waiting for the shot table
    if (it >= a.nt_min && it >= sh.nt) return;
    constexpr int RPT = TZ / BLOCK_Y;
    __shared__ T sP1[TZ + 1][TILE_X + 8], sP2[TZ + 1][TILE_X + 8], sP3x[TZ + 1][TILE_X + 8];
    __shared__ T sP3z[TZ + 8][TILE_X], sP4[TZ + 8][TILE_X], sP5[TZ + 8][TILE_X];
    constexpr int NK = FDC<FD>::NK;
    const int ld = g.ld, nz = g.nz, nx = g.nx;
    const int tz = (int)udiv((unsigned)bx, a.dtx), tx = bx - tz * a.tiles_x;
    const int r0 = tz * TZ, c0 = tx * TILE_X;
    const int lx = threadIdx.x, ty = threadIdx.y, tid = ty * TILE_X + lx;
    T *st = a.state + s * a.sstride + g.origin;
    T *__restrict__ U = st;
    T *__restrict__ V = st + g.fstride;
    const T *__restrict__ ex = st + 2 * g.fstride;
    const T *__restrict__ ez = st + 3 * g.fstride;
    const T *__restrict__ es = st + 4 * g.fstride;
    const T *__restrict__ L2M = a.L2M + g.origin;
    const T *__restrict__ M = a.M + g.origin;
    const T *__restrict__ B = a.B + g.origin;
    // stress products: (lambda+2mu) exx, lambda ezz, 2 mu exz along x;
    // 2 mu exz, (lambda+2mu) ezz, lambda exx along z.  Every load of the
    // stage (z-slab, x-halo, owned values) is issued before the first use:
    // cells outside the grid load cell 0 and are zeroed by a select.
    constexpr int KZ = (TZ + 8) / BLOCK_Y, KX = (8 * TZ + 255) / 256;
    const int c = c0 + lx;
    T zl[KZ], zm[KZ], zx[KZ], zz[KZ], zs[KZ];
    bool zok[KZ];
#pragma unroll
    for (int k = 0; k < KZ; ++k) {
        const int r = r0 - 4 + ty + BLOCK_Y * k;
        zok[k] = r >= 0 && r < nz && c < nx;
        const int p = zok[k] ? r * ld + c : 0;
        zl[k] = ldg(L2M, p);
        zm[k] = ldg(M, p);
        zx[k] = ldg(ex, p);
        zz[k] = ldg(ez, p);
        zs[k] = ldg(es, p);
    }
    T xl[KX], xm[KX], xx[KX], xz[KX], xs[KX];
    bool xok[KX];
#pragma unroll
    for (int k = 0; k < KX; ++k) {
        if (!halo_round(k, 8 * TZ)) continue;
        const int e = tid + 256 * k;
        const int zi = e >> 3, h = e & 7, xi = h < 4 ? h : h + TILE_X;
        const int r = r0 + zi, cc = c0 - 4 + xi;
        xok[k] = ((8 * TZ) % 256 == 0 || e < 8 * TZ) && r < nz && cc >= 0 && cc < nx;
        const int p = xok[k] ? r * ld + cc : 0;
        xl[k] = ldg(L2M, p);
        xm[k] = ldg(M, p);
        xx[k] = ldg(ex, p);
        xz[k] = ldg(ez, p);
        xs[k] = ldg(es, p);
    }
    // owned values and, in tiles touching a strip, the owned cells' C-PML
    // memories (x: Dx sxx, Dx sxz-adjacent terms 0..2; z: 0..2)
    T u0[RPT], v0[RPT], b0[RPT], mx[RPT][3], mz[RPT][3];
    int ixo[RPT], izo[RPT];
    const bool inx = interior_x(g, c0), inz = interior_z(g, r0, TZ);
#pragma unroll
    for (int rr = 0; rr < RPT; ++rr) {
        const int r = r0 + ty + BLOCK_Y * rr;
        const bool ok = r < nz && c < nx;
        const int q = ok ? r * ld + c : 0;
        u0[rr] = ldg((const T *)U, q);
        v0[rr] = ldg((const T *)V, q);
        b0[rr] = ldg(B, q);
        ixo[rr] = izo[rr] = -1;
#pragma unroll
        for (int j = 0; j < 3; ++j) mx[rr][j] = mz[rr][j] = T(0);
        if (!inx) {
            ixo[rr] = ok ? xstrip(g, r, c) : -1;
            const T *pb = a.psix + s * a.psix_stride;
#pragma unroll
            for (int j = 0; j < 3; ++j) mx[rr][j] = ldg(pb + j * a.psix_n, ixo[rr] >= 0 ? ixo[rr] : 0);
        }
        if (!inz) {
            izo[rr] = ok ? zstrip(g, r, c) : -1;
            const T *pb = a.psiz + s * a.psiz_stride;
#pragma unroll
            for (int j = 0; j < 3; ++j) mz[rr][j] = ldg(pb + j * a.psiz_n, izo[rr] >= 0 ? izo[rr] : 0);
        }
    }
#pragma unroll
    for (int k = 0; k < KZ; ++k) {
        const int zi = ty + BLOCK_Y * k;
        const T l2m = zl[k], mu = zm[k], exv = zx[k], ezv = zz[k];
        const T lam = l2m - T(2) * mu;
        const T p3 = zok[k] ? T(2) * mu * zs[k] : T(0);
        sP3z[zi][lx] = p3;
        sP4[zi][lx] = zok[k] ? l2m * ezv : T(0);
        sP5[zi][lx] = zok[k] ? lam * exv : T(0);
        if (k >= 1 && k < TZ / BLOCK_Y + 1) {
            sP1[zi - 4][lx + 4] = zok[k] ? l2m * exv : T(0);
            sP2[zi - 4][lx + 4] = zok[k] ? lam * ezv : T(0);
            sP3x[zi - 4][lx + 4] = p3;
        }
    }
#pragma unroll
    for (int k = 0; k < KX; ++k) {
        if (!halo_round(k, 8 * TZ)) continue;
        const int e = tid + 256 * k;
        {
            // lanes past the halo store into the spare row TZ
            const bool valid = (8 * TZ) % 256 == 0 || e < 8 * TZ;
            const int zi = valid ? e >> 3 : TZ, h = e & 7, xi = h < 4 ? h : h + TILE_X;
            const T l2m = xl[k], mu = xm[k];
            sP1[zi][xi] = xok[k] ? l2m * xx[k] : T(0);
            sP2[zi][xi] = xok[k] ? (l2m - T(2) * mu) * xz[k] : T(0);
            sP3x[zi][xi] = xok[k] ? T(2) * mu * xs[k] : T(0);
        }
    }
    __syncthreads();
    auto body = [&](auto tagx, auto tagz) {
        constexpr bool IX = decltype(tagx)::value, IZ = decltype(tagz)::value;
        // x-profile values of this lane's column (clamped into the profile)
        PmlFwd<T> kx{};
        if (!IX) kx = pml_fwd_coef(a.profx, c < nx ? c : nx - 1, nx);
#pragma unroll
        for (int rr = 0; rr < RPT; ++rr) {
            const int lz = ty + BLOCK_Y * rr, r = r0 + lz;
            if ((!IZ && r > nz - 2) || (!IX && (c < 1 || c > nx - 2))) continue;
            const int q = r * ld + c;
            const int xc = lx + 4, zc = lz + 4;
            T D1 = 0, D2 = 0, D3 = 0, D4 = 0, D5 = 0, D6 = 0;
#pragma unroll
            for (int k = 1; k <= NK; ++k) {
                const T ck = (T)fdc<FD>(k);
                // D+x, term range [k-1, nx-1-k]
                if (CX(c >= k - 1 && c <= nx - 1 - k)) {
                    D1 += ck * (sP1[lz][xc + k] - sP1[lz][xc - k + 1]);
                    D2 += ck * (sP2[lz][xc + k] - sP2[lz][xc - k + 1]);
                }
                // D-z of 2 mu exz, term 1 rows [1, nz-2], term k rows [k, nz-k]
                D3 += (CZ(k == 1 ? (r >= 1 && r <= nz - 2) : (r >= k && r <= nz - k)) ? ck : T(0)) *
                      (sP3z[zc + k - 1][lx] - sP3z[zc - k][lx]);
                // D-x of 2 mu exz, term 1 cols [1, nx-2], term k cols [k, nx-k]
                if (CX(k == 1 || (c >= k && c <= nx - k))) D4 += ck * (sP3x[lz][xc + k - 1] - sP3x[lz][xc - k]);
                // D+z, term range [k-1, nz-1-k]
                const T w56 = CZ(r >= k - 1 && r <= nz - 1 - k) ? ck : T(0);
                D5 += w56 * (sP4[zc + k][lx] - sP4[zc - k + 1][lx]);
                D6 += w56 * (sP5[zc + k][lx] - sP5[zc - k + 1][lx]);
            }
            if (!IZ && g.fs && r == 0) D3 = T(2 * fdc<FD>(1)) * sP3z[zc][lx];   // image of sigma_xz at the surface
            if (!IX) {
                const int ix = ixo[rr];
                if (ix >= 0) {
                    T *px = a.psix + s * a.psix_stride + ix;
                    D1 = pml_fwd_apply(D1, mx[rr][0], px + 0 * a.psix_n, kx);
                    D2 = pml_fwd_apply(D2, mx[rr][1], px + 1 * a.psix_n, kx);
                    D4 = pml_fwd_apply(D4, mx[rr][2], px + 2 * a.psix_n, kx);
                }
            }
            if (!IZ) {
                const int iz = izo[rr];
                if (iz >= 0) {
                    T *pz = a.psiz + s * a.psiz_stride + iz;
                    const PmlFwd<T> kz = pml_fwd_coef(a.profz, r, nz);
                    D3 = pml_fwd_apply_nt(D3, mz[rr][0], pz + 0 * a.psiz_n, kz);
                    D5 = pml_fwd_apply_nt(D5, mz[rr][1], pz + 1 * a.psiz_n, kz);
                    D6 = pml_fwd_apply_nt(D6, mz[rr][2], pz + 2 * a.psiz_n, kz);
                }
            }
            const T b = b0[rr];
            stg(U, q, u0[rr] + g.dtdx * b * D1 + g.dtdx * b * D2 + g.dtdz * b * D3);
            T v = v0[rr] + g.dtdx * b * D4 + g.dtdz * b * D5 + g.dtdz * b * D6;
            T sv = 0;
            if (sh.type == 0 && r == sh.sz && c == sh.sx) {
                const T w = a.wav[(i64)s * a.nt_max + it];
                v += g.fsrc * b * w;
                sv = g.fsrc * w;
            }
            stg(V, q, v);
            if (a.hist2_mode == 1) {
                const int rp = r - g.z0, cp = c - g.pad;
                if (CZ(rp >= 0 && rp < g.nzd) && CX(cp >= 0 && cp < g.nxd)) {
                    T *h = a.hist2 + ((i64)sh.hoff + (it - a.hist_t0)) * 2 * g.np;
                    const int o = rp * g.nxd + cp;
                    stst<T>(h, o, fma(g.dtdz, D3, fma(g.dtdx, D2, g.dtdx * D1)));
                    stst<T>(h, o + (int)g.np, fma(g.dtdz, D6, fma(g.dtdz, D5, g.dtdx * D4)) + sv);
                }
            }
        }
    };
    STFWI_DISPATCH(inx, inz, body);
}

// strain half-step (fwi/forward.py:369-444) + vx/vz receivers, history,
// illumination energy, snapshots
template <typename T, int FD, int TZ, bool LEAN>
__global__ __launch_bounds__(256) void k_fwd_strain(FwdArgs<T> a, int it)
{
    if constexpr (LEAN) {
        // no illumination energy, snapshots or velocity history
        a.illum = nullptr;
        a.snap = nullptr;
        a.hist2_mode = 0;
    }
    const Geo<T> &g = a.g;
    int bx, s;
    logical_block(a.dgx, bx, s);
    s += a.shot0;
    const ShotDev sh = a.shots[s];
    if (bx >= a.ntiles) {
        const int j = (bx - a.ntiles) * 256 + threadIdx.y * TILE_X + threadIdx.x;
        if (it >= sh.nt || j >= sh.nrec || !a.rec) return;
        const i64 cell = g.origin + a.rec_cell[sh.rec_off + j];
        const T *st = a.state + s * a.sstride;
        T *ro = a.rec + (i64)s * 4 * a.nt_max * a.nrec_max + (i64)it * a.nrec_max + j;
        if (a.rec_mask & 1) ro[0] = st[cell];
        if (a.rec_mask & 2) ro[(i64)a.nt_max * a.nrec_max] = st[g.fstride + cell];
        return;
    }
    // a shot past its record has no work at this step.  Only launches at or
    // past the shortest record among their shots test it: the others
    // issue their staging loads without waiting for the shot table
    if (it >= a.nt_min && it >= sh.nt) return;
    constexpr int RPT = TZ / BLOCK_Y;
    __shared__ T sU[TZ + 9][TILE_X + 8], sV[TZ + 9][TILE_X + 8];
    constexpr int NK = FDC<FD>::NK;
    const int ld = g.ld, nz = g.nz, nx = g.nx;
    const int tz = (int)udiv((unsigned)bx, a.dtx), tx = bx - tz * a.tiles_x;
    const int r0 = tz * TZ, c0 = tx * TILE_X;
    const int lx = threadIdx.x, ty = threadIdx.y, tid = ty * TILE_X + lx;
    T *st = a.state + s * a.sstride + g.origin;
    const T *__restrict__ U = st;
    const T *__restrict__ V = st + g.fstride;
    T *__restrict__ ex = st + 2 * g.fstride;
    T *__restrict__ ez = st + 3 * g.fstride;
    T *__restrict__ es = st + 4 * g.fstride;
    // velocities of the tile + 4-cell halo; all loads issued first, cells
    // outside the grid load cell 0 and are zeroed by a select
    constexpr int KZ = (TZ + 8) / BLOCK_Y, KX = (8 * (TZ + 8) + 255) / 256;
    const int c = c0 + lx;
    T zu[KZ], zv[KZ];
    bool zok[KZ];
#pragma unroll
    for (int k = 0; k < KZ; ++k) {
        const int r = r0 - 4 + ty + BLOCK_Y * k;
        zok[k] = r >= 0 && r < nz && c < nx;
        const int p = zok[k] ? r * ld + c : 0;
        zu[k] = ldgg(U, p);
        zv[k] = ldgg(V, p);
    }
    T xu[KX], xv[KX];
    bool xok[KX];
#pragma unroll
    for (int k = 0; k < KX; ++k) {
        if (!halo_round(k, 8 * (TZ + 8))) continue;
        const int e = tid + 256 * k;
        const int zi = e >> 3, h = e & 7, xi = h < 4 ? h : h + TILE_X;
        const int r = r0 - 4 + zi, cc = c0 - 4 + xi;
        xok[k] = ((8 * (TZ + 8)) % 256 == 0 || e < 8 * (TZ + 8)) && r >= 0 && r < nz && cc >= 0 && cc < nx;
        const int p = xok[k] ? r * ld + cc : 0;
        xu[k] = ldgg(U, p);
        xv[k] = ldgg(V, p);
    }
    // owned strains and, in tiles touching a strip, the owned cells' C-PML
    // memories 3 (exx / ezz derivatives) and 4 (exz derivatives)
    T ex0[RPT], ez0[RPT], es0[RPT], mx3[RPT], mx4[RPT], mz3[RPT], mz4[RPT];
    int ixo[RPT], izo[RPT];
    const bool inx = interior_x(g, c0), inz = interior_z(g, r0, TZ);
#pragma unroll
    for (int rr = 0; rr < RPT; ++rr) {
        const int r = r0 + ty + BLOCK_Y * rr;
        const bool ok = r < nz && c < nx;
        const int q = ok ? r * ld + c : 0;
        ex0[rr] = ldgg((const T *)ex, q);
        ez0[rr] = ldgg((const T *)ez, q);
        es0[rr] = ldgg((const T *)es, q);
        ixo[rr] = izo[rr] = -1;
        mx3[rr] = mx4[rr] = mz3[rr] = mz4[rr] = T(0);
        if (!inx) {
            ixo[rr] = ok ? xstrip(g, r, c) : -1;
            const T *pb = a.psix + s * a.psix_stride;
            const int i = ixo[rr] >= 0 ? ixo[rr] : 0;
            mx3[rr] = ldgg(pb + 3 * a.psix_n, i);
            mx4[rr] = ldgg(pb + 4 * a.psix_n, i);
        }
        if (!inz) {
            izo[rr] = ok ? zstrip(g, r, c) : -1;
            const T *pb = a.psiz + s * a.psiz_stride;
            const int i = izo[rr] >= 0 ? izo[rr] : 0;
            mz3[rr] = ldgg(pb + 3 * a.psiz_n, i);
            mz4[rr] = ldgg(pb + 4 * a.psiz_n, i);
        }
    }
#pragma unroll
    for (int k = 0; k < KZ; ++k) {
        const int zi = ty + BLOCK_Y * k;
        sU[zi][lx + 4] = zok[k] ? zu[k] : T(0);
        sV[zi][lx + 4] = zok[k] ? zv[k] : T(0);
    }
#pragma unroll
    for (int k = 0; k < KX; ++k) {
        if (!halo_round(k, 8 * (TZ + 8))) continue;
        const int e = tid + 256 * k;
        {
            // lanes past the halo store into the spare row TZ + 8
            const bool valid = (8 * (TZ + 8)) % 256 == 0 || e < 8 * (TZ + 8);
            const int zi = valid ? e >> 3 : TZ + 8, h = e & 7, xi = h < 4 ? h : h + TILE_X;
            sU[zi][xi] = xok[k] ? xu[k] : T(0);
            sV[zi][xi] = xok[k] ? xv[k] : T(0);
        }
    }
    __syncthreads();
    auto body = [&](auto tagx, auto tagz) {
        constexpr bool IX = decltype(tagx)::value, IZ = decltype(tagz)::value;
        // x-profile values of this lane's column (clamped into the profile)
        PmlFwd<T> kx{};
        if (!IX) kx = pml_fwd_coef(a.profx, c < nx ? c : nx - 1, nx);
#pragma unroll
        for (int rr = 0; rr < RPT; ++rr) {
            const int lz = ty + BLOCK_Y * rr, r = r0 + lz;
            if ((!IZ && r >= nz) || (!IX && c >= nx)) continue;
            const int xc = lx + 4, zc = lz + 4;
            const int q = r * ld + c;
            // exx: D-x U on columns [1, nx-1]
            T dxu = 0;
            if (CX(c >= 1)) {
#pragma unroll
                for (int k = 1; k <= NK; ++k)
                    if (CX(k == 1 || (c >= k && c <= nx - k)))
                        dxu += (T)fdc<FD>(k) * (sU[zc][xc + k - 1] - sU[zc][xc - k]);
            }
            const int ix = IX ? -1 : ixo[rr], iz = IZ ? -1 : izo[rr];
            T *px = a.psix + s * a.psix_stride + ix;
            T *pz = a.psiz + s * a.psiz_stride + iz;
            if (!IX && ix >= 0) dxu = pml_fwd_apply(dxu, mx3[rr], px + 3 * a.psix_n, kx);
            // z-profile values of this row, for D-z V and D+z U
            PmlFwd<T> kz{};
            if (!IZ && iz >= 0) kz = pml_fwd_coef(a.profz, r, nz);
            T exn = CX(c >= 1) ? ex0[rr] + g.dtdx * dxu : ex0[rr];
            // ezz: D-z V on columns [0, nx-2]; free-surface row from the C-PML'd D-x U
            T dzv = 0;
            if (CX(c <= nx - 2)) {
                if (!IZ && g.fs && r == 0) {
                    dzv = a.Lr0[c] * dxu * g.dz / g.dx;
                } else {
#pragma unroll
                    for (int k = 1; k <= NK; ++k)
                        dzv += (CZ(k == 1 ? r >= 1 : (r >= k && r <= nz - k)) ? (T)fdc<FD>(k) : T(0)) *
                               (sV[zc + k - 1][xc] - sV[zc - k][xc]);
                }
            }
            if (!IZ && iz >= 0) dzv = pml_fwd_apply_nt(dzv, mz3[rr], pz + 3 * a.psiz_n, kz);
            T ezn = ez0[rr] + g.dtdz * dzv;
            if (sh.type == 1 && r == sh.sz && c == sh.sx) {
                const T lm = (a.L2M[g.origin + q] - T(2) * a.M[g.origin + q]) + a.M[g.origin + q];
                const T stm = a.wav[(i64)s * a.nt_max + it] * g.dt / (T(2) * lm * g.dx * g.dz);
                exn += stm;
                ezn += stm;
            }
            stgg(ex, q, exn);
            stgg(ez, q, ezn);
            // exz on [0, nz-2] x [0, nx-2]
            T esn = es0[rr];
            if (CZ(r <= nz - 2) && CX(c <= nx - 2)) {
                T dzu = 0, dxv = 0;
#pragma unroll
                for (int k = 1; k <= NK; ++k) {
                    const T ck = (T)fdc<FD>(k);
                    // term masks as selects of the coefficient: no branches
                    dzu += (CX(c >= 1) && CZ(r >= k - 1 && r <= nz - 1 - k) ? ck : T(0)) *
                           (sU[zc + k][xc] - sU[zc - k + 1][xc]);
                    dxv += (CX(c >= k - 1 && c <= nx - 1 - k) ? ck : T(0)) * (sV[zc][xc + k] - sV[zc][xc - k + 1]);
                }
                if (!IZ && iz >= 0) dzu = pml_fwd_apply_nt(dzu, mz4[rr], pz + 4 * a.psiz_n, kz);
                if (!IX && ix >= 0) dxv = pml_fwd_apply(dxv, mx4[rr], px + 4 * a.psix_n, kx);
                esn = esn + T(0.5) * fma(g.dtdz, dzu, g.dtdx * dxv);
                stgg(es, q, esn);
            }
            if (!IZ && a.dxu0 && r == 0) a.dxu0[((i64)s * a.nt_max + it) * nx + c] = dxu;
            const int rp = r - g.z0, cp = c - g.pad;
            if (a.hist && CZ(rp >= 0 && rp < g.nzd) && CX(cp >= 0 && cp < g.nxd)) {
                T *hb = a.hist + ((i64)sh.hoff + (it - a.hist_t0)) * 3 * g.hnp;
                const int oh = rp * g.hpitch + cp + g.hcoff, nph = (int)g.hnp;
                stst(hb, oh, exn);
                stst(hb, oh + nph, ezn);
                stst(hb, oh + 2 * nph, esn);
                const int o = rp * g.nxd + cp, npi = (int)g.np;
                if (a.hist2_mode == 2) {
                    T *h = a.hist2 + ((i64)sh.hoff + (it - a.hist_t0)) * 2 * g.np;
                    stst(h, o, sU[zc][xc]);
                    stst(h, o + npi, sV[zc][xc]);
                }
            }
            const i64 pc = (i64)r * nx + c;
            if (a.illum) {
                const T f = a.illum_comp == 0 ? sU[zc][xc] : a.illum_comp == 1 ? sV[zc][xc] : a.illum_comp == 2 ? exn : ezn;
                a.illum[(i64)s * nz * nx + pc] += f * f;
            }
            if (a.snap) {
                const T vals[5] = {sU[zc][xc], sV[zc][xc], exn, ezn, esn};
                int k = 0;
                for (int f = 0; f < 5; ++f)
                    if (a.snap_mask & (1 << f)) {
                        stst(a.snap + (((i64)s * a.nsnap + k) * a.nt_max + it) * nz * nx, (int)pc, vals[f]);
                        ++k;
                    }
            }
        }
    };
    STFWI_DISPATCH(inx, inz, body);
}



// ---------------------------------------------------------------------------
// adjoint
// ---------------------------------------------------------------------------
template <typename T>
struct AdjArgs {
    Geo<T> g;
    const ShotDev *shots;
    const T *L2M, *M, *B, *Bimg, *Lr0;
    T *state;                          // per shot: U_a V_a exx_a ezz_a exz_a
    i64 sstride;
    const T *psix_o, *psiz_o;          // adjoint C-PML memories, previous step
    T *psix_n, *psiz_n;                // adjoint C-PML memories, this step
    i64 psix_stride, psiz_stride, psix_len, psiz_len;
    const T *profx, *profz;
    const T *wav;
    int nt_max;
    const T *rec_adj;                  // [shot][4][nt_max][nrec_max]
    int nrec_max;
    const i64 *uc_cell;                // merged receiver cells
    const int *uc_ptr, *uc_rec;        // CSR of receiver indices per merged cell
    const int *cellmap;                // [shot][fstride]: merged receiver index of a cell, -1 if none
    const int *rowcum;                 // [shot][nz + 1]: merged receiver cells in rows < r
    const T *injm;                     // [shot][nt_max][4][nuc_max]: summed receiver adjoints
    int nuc_max;
    T *src_adj;                        // [shot][nt_max]
    int imaging, gparams, nacc;
    const T *hist, *hist2;             // [frame][3][hnp] / [frame][2][np], shot s from frame hoff
    int hist_nt, hist_t0;              // frames per shot with checkpoints; step of frame 0
    const T *dxu0;
    T *gacc;                           // [shot][nacc][hnp]: rows of hpitch, as the history
    T *snap;
    int snap_mask, nsnap;
    T *aen;                            // [shot][nz*nx]: sum over steps of the squared adjoint field aen_comp
    int aen_comp;
    int tiles_x, tiles_z, ntiles;
    int shot0;                         // first shot of this launch's group
    int nt_min;                        // shortest record among the launch's shots
    UDiv dgx, dtx;                     // division by gridDim.x and by tiles_x
    int lean;                          // the LEAN instantiations may be used
};

// Receiver-adjoint injection (fwi/adjoint_jax_kernel.py:201-206; duplicate
// receivers accumulate).  The adjoints of each merged receiver cell are
// summed once per time loop into injm; the time-step kernels then add them:
// channels vx, vz to U_a, V_a as the strain-adjoint kernel loads its own
// cells, and channels exx, ezz of step it-1 to the strains the
// velocity-adjoint kernel of step it stores.  Only a shot's first reverse
// step (it = nt - 1) needs the separate launch below for exx, ezz.
template <typename T>
__global__ void k_merge_inj(AdjArgs<T> a, int nt_run)
{
    const int s = blockIdx.z;
    const ShotDev sh = a.shots[s];
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= sh.nuc) return;
    const int ui = sh.uc_off + u;
    const int p0 = a.uc_ptr[ui], p1 = a.uc_ptr[ui + 1];
    const int tn = sh.nt < nt_run ? sh.nt : nt_run;
    for (int t = blockIdx.y; t < tn; t += gridDim.y) {
        const T *ra = a.rec_adj + (i64)s * 4 * a.nt_max * a.nrec_max + (i64)t * a.nrec_max;
        T *o = (T *)a.injm + (((i64)s * a.nt_max + t) * 4) * a.nuc_max + u;
        for (int ch = 0; ch < 4; ++ch) {
            T sum = 0;
            for (int p = p0; p < p1; ++p) sum += ra[(i64)ch * a.nt_max * a.nrec_max + a.uc_rec[p]];
            o[(i64)ch * a.nuc_max] = sum;
        }
    }
}

// cellmap[s][origin + cell of merged receiver u] = u (the map is -1 elsewhere)
__global__ void k_cellmap(const ShotDev *shots, const i64 *uc_cell, i64 fstride, i64 origin, int *cellmap)
{
    const int s = blockIdx.y;
    const ShotDev sh = shots[s];
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= sh.nuc) return;
    cellmap[s * fstride + origin + uc_cell[sh.uc_off + u]] = u;
}

// exx, ezz receiver adjoints of a shot's first reverse step
template <typename T>
__global__ void k_adj_inject(AdjArgs<T> a, int it)
{
    const int s = a.shot0 + blockIdx.z;
    const ShotDev sh = a.shots[s];
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (it != sh.nt - 1 || u >= sh.nuc) return;
    const i64 cell = a.g.origin + a.uc_cell[sh.uc_off + u];
    T *st = a.state + s * a.sstride;
    const T *ij = a.injm + (((i64)s * a.nt_max + it) * 4) * a.nuc_max + u;
    for (int ch = 2; ch < 4; ++ch) st[ch * a.g.fstride + cell] += ij[(i64)ch * a.nuc_max];
}

// number of merged receiver cells in rows [r0, r0 + tz) of shot s
template <typename T>
__device__ __forceinline__ int tile_receivers(const AdjArgs<T> &a, int s, int r0, int tz)
{
    const int *rc = a.rowcum + (i64)s * (a.g.nz + 1);
    const int r1 = r0 + tz < a.g.nz ? r0 + tz : a.g.nz;
    return rc[r1] - rc[r0];
}

// adjoints of the four strain-update derivatives at any cell, after the
// transposed C-PML (fwi/adjoint_jax_kernel.py:212-301)
template <typename T>
struct K4Src {
    const Geo<T> &g;
    const T *esa, *exa, *eza;
    const T *pxo, *pzo;   // old psi base (per shot), arrays: x[0]=Dx_V x[1]=Dx_U, z[0]=Dz_U z[1]=Dz_V
    i64 pxn, pzn;
    const T *profx, *profz;
    const T *Lr0;
    __device__ __forceinline__ T GzU(int r, int c) const
    {
        return (r <= g.nz - 2 && c <= g.nx - 2) ? g.hdtdz * esa[r * g.ld + c] : T(0);
    }
    __device__ __forceinline__ T GxV(int r, int c) const
    {
        return (r <= g.nz - 2 && c <= g.nx - 2) ? g.hdtdx * esa[r * g.ld + c] : T(0);
    }
    __device__ __forceinline__ T GzV(int r, int c) const { return g.dtdz * eza[r * g.ld + c]; }
    __device__ __forceinline__ T DzVa(int r, int c) const
    {
        const T G = GzV(r, c);
        const int iz = zstrip(g, r, c);
        return iz >= 0 ? pmlT_val(G, pzo[pzn + iz], profz, r, g.nz) : G;
    }
    __device__ __forceinline__ T GxU(int r, int c) const
    {
        T G = 0;
        if (g.fs && r == 0 && c <= g.nx - 2) G = Lr0[c] * g.dzdx * DzVa(0, c);
        if (c >= 1) G = G + g.dtdx * exa[r * g.ld + c];
        return G;
    }
    __device__ __forceinline__ T DzUa(int r, int c) const
    {
        const T G = GzU(r, c);
        const int iz = zstrip(g, r, c);
        return iz >= 0 ? pmlT_val(G, pzo[iz], profz, r, g.nz) : G;
    }
    __device__ __forceinline__ T DxVa(int r, int c) const
    {
        const T G = GxV(r, c);
        const int ix = xstrip(g, r, c);
        return ix >= 0 ? pmlT_val(G, pxo[ix], profx, c, g.nx) : G;
    }
    __device__ __forceinline__ T DxUa(int r, int c) const
    {
        const T G = GxU(r, c);
        const int ix = xstrip(g, r, c);
        return ix >= 0 ? pmlT_val(G, pxo[pxn + ix], profx, c, g.nx) : G;
    }
};

// strain-adjoint sweep: steps 0-T..2-T of fwi/adjoint_jax_kernel.py:201-323;
// produces the final U_a, V_a of the step and the source-wavelet adjoint
template <typename T, int FD, int TZ, bool LEAN>
__global__ __launch_bounds__(256) void k_adj_strain(AdjArgs<T> a, int it)
{
    if constexpr (LEAN) {
        // no source-wavelet adjoint
        a.src_adj = nullptr;
    }
    const Geo<T> &g = a.g;
    int bx, s;
    logical_block(a.dgx, bx, s);
    s += a.shot0;
    const ShotDev sh = a.shots[s];
    // a shot past its record has no work at this step.  Only launches at or
    // past the shortest record among their shots test it: the others
    // issue their staging loads without waiting for the shot table
    if (it >= a.nt_min && it >= sh.nt) return;
    constexpr int RPT = TZ / BLOCK_Y;
    __shared__ T szU[TZ + 8][TILE_X], szV[TZ + 8][TILE_X];
    __shared__ T sxV[TZ + 1][TILE_X + 8], sxU[TZ + 1][TILE_X + 8];
    constexpr int NK = FDC<FD>::NK;
    const int nz = g.nz, nx = g.nx, ld = g.ld;
    const int tz = (int)udiv((unsigned)bx, a.dtx), tx = bx - tz * a.tiles_x;
    const int r0 = tz * TZ, c0 = tx * TILE_X;
    const int lx = threadIdx.x, ty = threadIdx.y, tid = ty * TILE_X + lx;
    T *st = a.state + s * a.sstride + g.origin;
    T *Ua = st, *Va = st + g.fstride;
    const K4Src<T> S{g, st + 4 * g.fstride, st + 2 * g.fstride, st + 3 * g.fstride,
                     a.psix_o + s * a.psix_stride, a.psiz_o + s * a.psiz_stride,
                     a.psix_len, a.psiz_len, a.profx, a.profz, a.Lr0};
    const bool inx = interior_x(g, c0), inz = interior_z(g, r0, TZ);
    const int c = c0 + lx;
    // owned values, loaded ahead of the barrier
    T ua0[RPT], va0[RPT];
#pragma unroll
    for (int rr = 0; rr < RPT; ++rr) {
        const int r = r0 + ty + BLOCK_Y * rr;
        const int q = (r < nz && c < nx) ? r * ld + c : 0;
        ua0[rr] = ldg((const T *)Ua, q);
        va0[rr] = ldg((const T *)Va, q);
    }
    // derivative adjoints after the transposed C-PML, tile + halo.  All
    // loads of the stage are issued before the first use; cells outside the
    // grid load cell 0 and are zeroed by a select.
    constexpr int KZ = (TZ + 8) / BLOCK_Y, KX = (8 * TZ + 255) / 256;
    const T *esa = S.esa, *exa = S.exa, *eza = S.eza;
    T zs[KZ], zx[KZ], zz[KZ];
    bool zok[KZ];
#pragma unroll
    for (int k = 0; k < KZ; ++k) {
        const int r = r0 - 4 + ty + BLOCK_Y * k;
        zok[k] = r >= 0 && r < nz && c < nx;
        const int p = zok[k] ? r * ld + c : 0;
        zs[k] = ldg(esa, p);
        zx[k] = ldg(exa, p);
        zz[k] = ldg(eza, p);
    }
    T hs[KX], hx[KX], hz[KX];
    bool hok[KX];
#pragma unroll
    for (int k = 0; k < KX; ++k) {
        if (!halo_round(k, 8 * TZ)) continue;
        const int e = tid + 256 * k;
        const int zi = e >> 3, h = e & 7, xi = h < 4 ? h : h + TILE_X;
        const int r = r0 + zi, cc = c0 - 4 + xi;
        hok[k] = ((8 * TZ) % 256 == 0 || e < 8 * TZ) && r < nz && cc >= 0 && cc < nx;
        const int p = hok[k] ? r * ld + cc : 0;
        hs[k] = ldg(esa, p);
        hx[k] = ldg(exa, p);
        hz[k] = ldg(eza, p);
    }
    // C-PML inputs of the staged cells and the owned cells' G values, kept
    // for the memory update after the barrier (tiles touching a strip only)
    PmlIn<T> pz[KZ] = {}, px[KZ] = {};
    int izs[KZ], ixs[KZ];
    T gzu_o[RPT], gzv_o[RPT], gxv_o[RPT], gxu_o[RPT];
#pragma unroll
    for (int k = 0; k < KZ; ++k) izs[k] = ixs[k] = -1;
#pragma unroll
    for (int rr = 0; rr < RPT; ++rr) gzu_o[rr] = gzv_o[rr] = gxv_o[rr] = gxu_o[rr] = T(0);
    // staging, compiled per tile class: along an interior axis the C-PML
    // inputs and the edge masks of the staged cells are compiled out
    auto stage = [&](auto tagx, auto tagz) {
        constexpr bool IX = decltype(tagx)::value, IZ = decltype(tagz)::value;
    if (IX && IZ) {
        // no PML, no masks: the adjoints are the scaled strain adjoints
#pragma unroll
        for (int k = 0; k < KZ; ++k) {
            const int zi = ty + BLOCK_Y * k;
            szU[zi][lx] = g.hdtdz * zs[k];
            szV[zi][lx] = g.dtdz * zz[k];
            if (k >= 1 && k < TZ / BLOCK_Y + 1) {
                sxV[zi - 4][lx + 4] = g.hdtdx * zs[k];
                sxU[zi - 4][lx + 4] = g.dtdx * zx[k];
            }
        }
#pragma unroll
        for (int k = 0; k < KX; ++k) {
            if (!halo_round(k, 8 * TZ)) continue;
            const int e = tid + 256 * k;
            {
                // lanes past the halo store into the spare row TZ
                const bool valid = (8 * TZ) % 256 == 0 || e < 8 * TZ;
                const int zi = valid ? e >> 3 : TZ, h = e & 7, xi = h < 4 ? h : h + TILE_X;
                sxV[zi][xi] = g.hdtdx * hs[k];
                sxU[zi][xi] = g.dtdx * hx[k];
            }
        }
    } else {
        // z-slab: Dz_U_a, Dz_V_a everywhere, Dx_V_a, Dx_U_a on the tile rows
        PmlIn<T> kxc{};   // x-profile values of this lane's column
        if (!IX) kxc.template load_coef<true>(a.profx, c, nx);
#pragma unroll
        for (int k = 0; k < KZ; ++k) {
            const int r = r0 - 4 + ty + BLOCK_Y * k;
            izs[k] = (!IZ && zok[k]) ? zstrip(g, r, c) : -1;
            ixs[k] = (!IX && zok[k]) ? xstrip(g, r, c) : -1;
            if (!IZ) pz[k].template load<true>(S.pzo, S.pzo + S.pzn, izs[k], a.profz, r, nz);
            if (!IX && k >= 1 && k < TZ / BLOCK_Y + 1)
                px[k].template load_mem<true>(S.pxo, S.pxo + S.pxn, ixs[k], kxc);
        }
        PmlIn<T> hp[KX];
        int ihs[KX];
#pragma unroll
        for (int k = 0; k < KX; ++k) {
            if (!halo_round(k, 8 * TZ)) continue;
            const int e = tid + 256 * k;
            const int zi = e >> 3, h = e & 7, xi = h < 4 ? h : h + TILE_X;
            ihs[k] = (!IX && hok[k]) ? xstrip(g, r0 + zi, c0 - 4 + xi) : -1;
            if (!IX) hp[k].template load<true>(S.pxo, S.pxo + S.pxn, ihs[k], a.profx, c0 - 4 + xi, nx);
        }
        // Dx_U_a input: dt/dx exx_a (columns >= 1) plus, on the free-surface
        // row, the image term L_ratio dz/dx Dz_V_a (row 0 has no z-strip)
        auto gxu = [&](int r, int cc, T exv, T ezv) {
            T G = 0;
            if (!IZ && g.fs && r == 0 && CX(cc <= nx - 2)) G = a.Lr0[cc] * g.dzdx * (g.dtdz * ezv);
            if (CX(cc >= 1)) G = G + g.dtdx * exv;
            return G;
        };
#pragma unroll
        for (int k = 0; k < KZ; ++k) {
            const int zi = ty + BLOCK_Y * k, r = r0 - 4 + zi;
            T u = 0, v = 0, xv = 0, xu = 0;
            if (zok[k]) {
                const bool in2 = CZ(r <= nz - 2) && CX(c <= nx - 2);
                const T gzu = in2 ? g.hdtdz * zs[k] : T(0);
                const T gzv = g.dtdz * zz[k];
                u = izs[k] >= 0 ? pmlT_pre(gzu, pz[k].m0, pz[k].q, pz[k].a) : gzu;
                v = izs[k] >= 0 ? pmlT_pre(gzv, pz[k].m1, pz[k].q, pz[k].a) : gzv;
                if (k >= 1 && k < TZ / BLOCK_Y + 1) {
                    const T gxv = in2 ? g.hdtdx * zs[k] : T(0);
                    const T gxu_ = gxu(r, c, zx[k], zz[k]);
                    xv = ixs[k] >= 0 ? pmlT_pre(gxv, px[k].m0, px[k].q, px[k].a) : gxv;
                    xu = ixs[k] >= 0 ? pmlT_pre(gxu_, px[k].m1, px[k].q, px[k].a) : gxu_;
                    gzu_o[k - 1] = gzu;
                    gzv_o[k - 1] = gzv;
                    gxv_o[k - 1] = gxv;
                    gxu_o[k - 1] = gxu_;
                }
            }
            szU[zi][lx] = u;
            szV[zi][lx] = v;
            if (k >= 1 && k < TZ / BLOCK_Y + 1) {
                sxV[zi - 4][lx + 4] = xv;
                sxU[zi - 4][lx + 4] = xu;
            }
        }
#pragma unroll
        for (int k = 0; k < KX; ++k) {
            if (!halo_round(k, 8 * TZ)) continue;
            const int e = tid + 256 * k;
            {
                // lanes past the halo store into the spare row TZ
                const bool valid = (8 * TZ) % 256 == 0 || e < 8 * TZ;
                const int zi = valid ? e >> 3 : TZ, h = e & 7, xi = h < 4 ? h : h + TILE_X;
                const int r = r0 + zi, cc = c0 - 4 + xi;
                T xv = 0, xu = 0;
                if (hok[k]) {
                    const T gxv = (CZ(r <= nz - 2) && CX(cc <= nx - 2)) ? g.hdtdx * hs[k] : T(0);
                    const T gxu_ = gxu(r, cc, hx[k], hz[k]);
                    xv = ihs[k] >= 0 ? pmlT_pre(gxv, hp[k].m0, hp[k].q, hp[k].a) : gxv;
                    xu = ihs[k] >= 0 ? pmlT_pre(gxu_, hp[k].m1, hp[k].q, hp[k].a) : gxu_;
                }
                sxV[zi][xi] = xv;
                sxU[zi][xi] = xu;
            }
        }
    }
    };
    STFWI_DISPATCH(inx, inz, stage);
    // receiver adjoints of channels vx, vz of this step into the owned
    // U_a, V_a (tiles holding receiver cells only)
    if (tile_receivers(a, s, r0, TZ) > 0) {
        const int *cm = a.cellmap + (i64)s * g.fstride + g.origin;
        const T *ij = a.injm + (((i64)s * a.nt_max + it) * 4) * a.nuc_max;
#pragma unroll
        for (int rr = 0; rr < RPT; ++rr) {
            const int r = r0 + ty + BLOCK_Y * rr;
            if (r < nz && c < nx) {
                const int u = cm[r * ld + c];
                if (u >= 0) {
                    ua0[rr] += ij[u];
                    va0[rr] += ij[a.nuc_max + u];
                }
            }
        }
    }
    __syncthreads();
    auto body = [&](auto tagx, auto tagz) {
        constexpr bool IX = decltype(tagx)::value, IZ = decltype(tagz)::value;
        // the column (x) term masks as per-lane weights, set once per tile:
        // a masked term adds 0 * value (the staged values are finite), an
        // unmasked one its coefficient; along an interior axis they fold to
        // the constant coefficients
        T wzu[NK], wvx1[NK], wvx2[NK], wzv[NK], wux1[NK], wux2[NK];
#pragma unroll
        for (int k = 1; k <= NK; ++k) {
            const T ck = (T)fdc<FD>(k);
            const int i1 = c - k + 1, i2 = c + k;
            wzu[k - 1] = CX(c >= 1) ? ck : T(0);
            wvx1[k - 1] = CX(c - k >= k - 1 && c - k <= nx - 1 - k) ? ck : T(0);
            wvx2[k - 1] = CX(c + k - 1 >= k - 1 && c + k - 1 <= nx - 1 - k) ? -ck : T(0);
            wzv[k - 1] = CX(c <= nx - 2) ? ck : T(0);
            wux1[k - 1] = CX(k == 1 ? (i1 >= 1 && i1 <= nx - 1) : (i1 >= k && i1 <= nx - k)) ? ck : T(0);
            wux2[k - 1] = CX(k == 1 ? (i2 >= 1 && i2 <= nx - 1) : (i2 >= k && i2 <= nx - k)) ? -ck : T(0);
        }
#pragma unroll
        for (int rr = 0; rr < RPT; ++rr) {
            const int lz = ty + BLOCK_Y * rr, r = r0 + lz;
            if ((!IZ && r >= nz) || (!IX && c >= nx)) continue;
            const int xc = lx + 4, zc = lz + 4;
            const int q = r * ld + c;
            T ua = ua0[rr], va = va0[rr];
            // the row (z) term masks as selects of the weights: no branches
            const bool rz = CZ(r <= nz - 2);
#pragma unroll
            for (int k = 1; k <= NK; ++k) {
                // (D+z)^T of Dz_U_a, columns [1, nx-1]; term range [k-1, nz-1-k]
                const T u1 = CZ(r - k >= k - 1 && r - k <= nz - 1 - k) ? wzu[k - 1] : T(0);
                const T u2 = CZ(r + k - 1 >= k - 1 && r + k - 1 <= nz - 1 - k) ? -wzu[k - 1] : T(0);
                ua += u1 * szU[zc - k][lx];
                ua += u2 * szU[zc + k - 1][lx];
                // (D+x)^T of Dx_V_a, rows [0, nz-2]
                va += (rz ? wvx1[k - 1] : T(0)) * sxV[lz][xc - k];
                va += (rz ? wvx2[k - 1] : T(0)) * sxV[lz][xc + k - 1];
            }
#pragma unroll
            for (int k = 1; k <= NK; ++k) {
                // (D-z)^T of Dz_V_a, columns [0, nx-2]; term 1 rows [1, nz-1], k: [k, nz-k]
                const int i1 = r - k + 1, i2 = r + k;
                const T v1 = CZ(k == 1 ? (i1 >= 1 && i1 <= nz - 1) : (i1 >= k && i1 <= nz - k)) ? wzv[k - 1] : T(0);
                const T v2 = CZ(k == 1 ? (i2 >= 1 && i2 <= nz - 1) : (i2 >= k && i2 <= nz - k)) ? -wzv[k - 1] : T(0);
                va += v1 * szV[zc - k + 1][lx];
                va += v2 * szV[zc + k][lx];
            }
#pragma unroll
            for (int k = 1; k <= NK; ++k) {
                // (D-x)^T of Dx_U_a, all rows; term 1 cols [1, nx-1], k: [k, nx-k]
                ua += wux1[k - 1] * sxU[lz][xc - k + 1];
                ua += wux2[k - 1] * sxU[lz][xc + k];
            }
            stg(Ua, q, ua);
            stg(Va, q, va);
            // C-PML memories of this cell: psi' = b (G + psi), from the values
            // staged for row rr + 1 of the z-slab
            if (!IZ) {
                const int iz = izs[rr + 1];
                if (iz >= 0) {
                    T *pzn = a.psiz_n + s * a.psiz_stride + iz;
                    const PmlIn<T> &m = pz[rr + 1];
                    __builtin_nontemporal_store(m.b * (gzu_o[rr] + m.m0), pzn + 0);
                    __builtin_nontemporal_store(m.b * (gzv_o[rr] + m.m1), pzn + a.psiz_len);
                }
            }
            if (!IX) {
                const int ix = ixs[rr + 1];
                if (ix >= 0) {
                    T *pxn = a.psix_n + s * a.psix_stride + ix;
                    const PmlIn<T> &m = px[rr + 1];
                    pxn[0] = m.b * (gxv_o[rr] + m.m0);
                    pxn[a.psix_len] = m.b * (gxu_o[rr] + m.m1);
                }
            }
            if (r == sh.sz && c == sh.sx && a.src_adj) {
                T sa;
                if (sh.type == 1) {
                    const T lm = (a.L2M[g.origin + q] - T(2) * a.M[g.origin + q]) + a.M[g.origin + q];
                    sa = (g.dt / (T(2) * lm * g.dx * g.dz)) * (S.exa[q] + S.eza[q]);
                } else {
                    sa = g.fsrc * a.B[g.origin + q] * va;
                }
                a.src_adj[(i64)s * a.nt_max + it] = sa;
            }
        }
    };
    STFWI_DISPATCH(inx, inz, body);
}


// adjoints of the velocity-update derivatives (fwi/adjoint_jax_kernel.py:325-399)
template <typename T>
struct K5Src {
    const Geo<T> &g;
    const T *Ua, *Va, *B;
    const T *pxo, *pzo;   // x[2]=Dx_M_es x[3]=Dx_(B U);  z[2]=Dz_(B V) z[3]=Dz_M_es
    i64 pxn, pzn;
    const T *profx, *profz;
    __device__ __forceinline__ bool inR(int r, int c) const { return r <= g.nz - 2 && c >= 1 && c <= g.nx - 2; }
    __device__ __forceinline__ T GxBU(int r, int c) const
    {
        const int p = r * g.ld + c;
        return inR(r, c) ? g.dtdx * B[p] * Ua[p] : T(0);
    }
    __device__ __forceinline__ T GzMes(int r, int c) const
    {
        const int p = r * g.ld + c;
        return inR(r, c) ? g.dtdz * B[p] * Ua[p] : T(0);
    }
    __device__ __forceinline__ T GxMes(int r, int c) const
    {
        const int p = r * g.ld + c;
        return inR(r, c) ? g.dtdx * B[p] * Va[p] : T(0);
    }
    __device__ __forceinline__ T GzBV(int r, int c) const
    {
        const int p = r * g.ld + c;
        return inR(r, c) ? g.dtdz * B[p] * Va[p] : T(0);
    }
    __device__ __forceinline__ T DxBUa(int r, int c) const
    {
        const T G = GxBU(r, c);
        const int ix = xstrip(g, r, c);
        return ix >= 0 ? pmlT_val(G, pxo[3 * pxn + ix], profx, c, g.nx) : G;
    }
    __device__ __forceinline__ T DxMesa(int r, int c) const
    {
        const T G = GxMes(r, c);
        const int ix = xstrip(g, r, c);
        return ix >= 0 ? pmlT_val(G, pxo[2 * pxn + ix], profx, c, g.nx) : G;
    }
    __device__ __forceinline__ T DzBVa(int r, int c) const
    {
        const T G = GzBV(r, c);
        const int iz = zstrip(g, r, c);
        return iz >= 0 ? pmlT_val(G, pzo[2 * pzn + iz], profz, r, g.nz) : G;
    }
    __device__ __forceinline__ T DzMesa(int r, int c) const
    {
        const T G = GzMes(r, c);
        const int iz = zstrip(g, r, c);
        return iz >= 0 ? pmlT_val(G, pzo[3 * pzn + iz], profz, r, g.nz) : G;
    }
};

// velocity-adjoint sweep with the imaging condition fused: steps 2-T (V),
// 1-T (U) and the constitutive transposes (fwi/adjoint_jax_kernel.py:325-464)
template <typename T, int FD, int TZ, bool LEAN>
__global__ __launch_bounds__(256) void k_adj_vel(AdjArgs<T> a, int it)
{
    if constexpr (LEAN) {
        // exact imaging of the Vs gradient only; no snapshots or energy
        a.snap = nullptr;
        a.aen = nullptr;
        a.imaging = 2;
        a.gparams = 1;
    }
    const Geo<T> &g = a.g;
    int bx, s;
    logical_block(a.dgx, bx, s);
    s += a.shot0;
    const ShotDev sh = a.shots[s];
    // a shot past its record has no work at this step.  Only launches at or
    // past the shortest record among their shots test it: the others
    // issue their staging loads without waiting for the shot table
    if (it >= a.nt_min && it >= sh.nt) return;
    constexpr int RPT = TZ / BLOCK_Y;
    __shared__ T sxBU[TZ + 1][TILE_X + 8], sxMes[TZ + 1][TILE_X + 8];
    __shared__ T szBV[TZ + 8][TILE_X], szMes[TZ + 8][TILE_X];
    constexpr int NK = FDC<FD>::NK;
    const int nz = g.nz, nx = g.nx, ld = g.ld;
    const int tz = (int)udiv((unsigned)bx, a.dtx), tx = bx - tz * a.tiles_x;
    const int r0 = tz * TZ, c0 = tx * TILE_X;
    // the row index is uniform across a wavefront (64 x 4 threads): made
    // scalar here, the z-dependent masks and C-PML row lookups of this kernel
    // become scalar work (measured faster for this kernel only)
    const int lx = threadIdx.x, ty = __builtin_amdgcn_readfirstlane(threadIdx.y), tid = ty * TILE_X + lx;
    T *st = a.state + s * a.sstride + g.origin;
    const T *Ua = st, *Va = st + g.fstride;
    T *exa = st + 2 * g.fstride, *eza = st + 3 * g.fstride, *esa = st + 4 * g.fstride;
    const T *Bp = a.B + g.origin;
    const K5Src<T> S{g, Ua, Va, Bp, a.psix_o + s * a.psix_stride, a.psiz_o + s * a.psiz_stride,
                     a.psix_len, a.psiz_len, a.profx, a.profz};
    const bool inx = interior_x(g, c0), inz = interior_z(g, r0, TZ);
    const int c = c0 + lx;
    // B, U_a, V_a of the tile + halo, then the owned values and the history
    // frame: all loads of the stage are issued before the first use; cells
    // outside the grid load cell 0 and are zeroed by a select
    constexpr int KZ = (TZ + 8) / BLOCK_Y, KX = (8 * TZ + 255) / 256;
    T zb[KZ], zu[KZ], zv[KZ];
    bool zok[KZ];
#pragma unroll
    for (int k = 0; k < KZ; ++k) {
        const int r = r0 - 4 + ty + BLOCK_Y * k;
        zok[k] = r >= 0 && r < nz && c < nx;
        const int p = zok[k] ? r * ld + c : 0;
        zb[k] = ldgg(Bp, p);
        zu[k] = ldgg(Ua, p);
        zv[k] = ldgg(Va, p);
    }
    T hb[KX], hu[KX], hv[KX];
    bool hok[KX];
#pragma unroll
    for (int k = 0; k < KX; ++k) {
        if (!halo_round(k, 8 * TZ)) continue;
        const int e = tid + 256 * k;
        const int zi = e >> 3, h = e & 7, xi = h < 4 ? h : h + TILE_X;
        const int r = r0 + zi, cc = c0 - 4 + xi;
        hok[k] = ((8 * TZ) % 256 == 0 || e < 8 * TZ) && r < nz && cc >= 0 && cc < nx;
        const int p = hok[k] ? r * ld + cc : 0;
        hb[k] = ldgg(Bp, p);
        hu[k] = ldgg(Ua, p);
        hv[k] = ldgg(Va, p);
    }
    // owned values, loaded ahead of the barrier
    T l2m0[RPT], mu0[RPT], exa0v[RPT], eza0v[RPT], esa0v[RPT], h0[RPT], h1[RPT], h2v[RPT], ac0[RPT], ac1[RPT];
    const i64 hs = (i64)s * a.nt_max;            // wavelet / dxu0 rows
    const i64 hh = (i64)sh.hoff - a.hist_t0;      // history frame base
    const int hframe = a.imaging == 2 ? it - 1 : it;
#pragma unroll
    for (int rr = 0; rr < RPT; ++rr) {
        const int r = r0 + ty + BLOCK_Y * rr;
        const int q = (r < nz && c < nx) ? r * ld + c : 0;
        l2m0[rr] = ldgg(a.L2M + g.origin, q);
        mu0[rr] = ldgg(a.M + g.origin, q);
        exa0v[rr] = ldgg((const T *)exa, q);
        eza0v[rr] = ldgg((const T *)eza, q);
        esa0v[rr] = ldgg((const T *)esa, q);
    }
    // receiver adjoints of channels exx, ezz of step it-1, added to the
    // strains this step stores (receiver tiles only)
    T inj2[RPT], inj3[RPT];
    bool hasinj[RPT];
#pragma unroll
    for (int rr = 0; rr < RPT; ++rr) {
        inj2[rr] = inj3[rr] = T(0);
        hasinj[rr] = false;
    }
    if (it >= 1 && tile_receivers(a, s, r0, TZ) > 0) {
        const int *cm = a.cellmap + (i64)s * g.fstride + g.origin;
        const T *ij = a.injm + (((i64)s * a.nt_max + it - 1) * 4 + 2) * a.nuc_max;
#pragma unroll
        for (int rr = 0; rr < RPT; ++rr) {
            const int r = r0 + ty + BLOCK_Y * rr;
            if (r < nz && c < nx) {
                const int u = cm[r * ld + c];
                if (u >= 0) {
                    hasinj[rr] = true;
                    inj2[rr] = ij[u];
                    inj3[rr] = ij[a.nuc_max + u];
                }
            }
        }
    }
    // C-PML inputs of the staged cells and the owned cells' G values, kept
    // for the memory update after the barrier (tiles touching a strip only)
    PmlIn<T> pz[KZ] = {}, px[KZ] = {};
    int izs[KZ], ixs[KZ];
    T gzbv_o[RPT], gzmes_o[RPT], gxbu_o[RPT], gxmes_o[RPT];
#pragma unroll
    for (int k = 0; k < KZ; ++k) izs[k] = ixs[k] = -1;
#pragma unroll
    for (int rr = 0; rr < RPT; ++rr) gzbv_o[rr] = gzmes_o[rr] = gxbu_o[rr] = gxmes_o[rr] = T(0);
    // staging, compiled per tile class: along an interior axis the C-PML
    // inputs and the edge masks of the staged cells are compiled out
    auto stage = [&](auto tagx, auto tagz) {
        constexpr bool IX = decltype(tagx)::value, IZ = decltype(tagz)::value;
    if (IX && IZ) {
        // no PML, no masks: Dx_(BU)_a = dt/dx B U_a, Dz_M_es_a = dt/dz B U_a, ...
#pragma unroll
        for (int k = 0; k < KZ; ++k) {
            const int zi = ty + BLOCK_Y * k;
            const T b = zb[k], u = zu[k], v = zv[k];
            szBV[zi][lx] = g.dtdz * b * v;
            szMes[zi][lx] = g.dtdz * b * u;
            if (k >= 1 && k < TZ / BLOCK_Y + 1) {
                sxBU[zi - 4][lx + 4] = g.dtdx * b * u;
                sxMes[zi - 4][lx + 4] = g.dtdx * b * v;
            }
        }
#pragma unroll
        for (int k = 0; k < KX; ++k) {
            if (!halo_round(k, 8 * TZ)) continue;
            const int e = tid + 256 * k;
            {
                // lanes past the halo store into the spare row TZ
                const bool valid = (8 * TZ) % 256 == 0 || e < 8 * TZ;
                const int zi = valid ? e >> 3 : TZ, h = e & 7, xi = h < 4 ? h : h + TILE_X;
                const T b = hb[k];
                sxBU[zi][xi] = g.dtdx * b * hu[k];
                sxMes[zi][xi] = g.dtdx * b * hv[k];
            }
        }
    } else {
        // memories: x[2] = Dx_M_es, x[3] = Dx_(B U); z[2] = Dz_(B V), z[3] = Dz_M_es
        PmlIn<T> kxc{};   // x-profile values of this lane's column
        if (!IX) kxc.template load_coef<false>(a.profx, c, nx);
#pragma unroll
        for (int k = 0; k < KZ; ++k) {
            const int r = r0 - 4 + ty + BLOCK_Y * k;
            izs[k] = (!IZ && zok[k]) ? zstrip(g, r, c) : -1;
            ixs[k] = (!IX && zok[k]) ? xstrip(g, r, c) : -1;
            if (!IZ) pz[k].template load<false>(S.pzo + 2 * S.pzn, S.pzo + 3 * S.pzn, izs[k], a.profz, r, nz);
            if (!IX && k >= 1 && k < TZ / BLOCK_Y + 1)
                px[k].template load_mem<false>(S.pxo + 2 * S.pxn, S.pxo + 3 * S.pxn, ixs[k], kxc);
        }
        PmlIn<T> hp[KX];
        int ihs[KX];
#pragma unroll
        for (int k = 0; k < KX; ++k) {
            if (!halo_round(k, 8 * TZ)) continue;
            const int e = tid + 256 * k;
            const int zi = e >> 3, h = e & 7, xi = h < 4 ? h : h + TILE_X;
            ihs[k] = (!IX && hok[k]) ? xstrip(g, r0 + zi, c0 - 4 + xi) : -1;
            if (!IX) hp[k].template load<false>(S.pxo + 2 * S.pxn, S.pxo + 3 * S.pxn, ihs[k], a.profx, c0 - 4 + xi, nx);
        }
#pragma unroll
        for (int k = 0; k < KZ; ++k) {
            const int zi = ty + BLOCK_Y * k, r = r0 - 4 + zi;
            T bv = 0, mz = 0, bu = 0, mx = 0;
            if (zok[k]) {
                const bool inR = CZ(r <= nz - 2) && CX(c >= 1 && c <= nx - 2);
                const T b = zb[k], u = zu[k], v = zv[k];
                const T gzbv = inR ? g.dtdz * b * v : T(0);
                const T gzmes = inR ? g.dtdz * b * u : T(0);
                bv = izs[k] >= 0 ? pmlT_pre(gzbv, pz[k].m0, pz[k].q, pz[k].a) : gzbv;
                mz = izs[k] >= 0 ? pmlT_pre(gzmes, pz[k].m1, pz[k].q, pz[k].a) : gzmes;
                if (k >= 1 && k < TZ / BLOCK_Y + 1) {
                    const T gxbu = inR ? g.dtdx * b * u : T(0);
                    const T gxmes = inR ? g.dtdx * b * v : T(0);
                    bu = ixs[k] >= 0 ? pmlT_pre(gxbu, px[k].m1, px[k].q, px[k].a) : gxbu;
                    mx = ixs[k] >= 0 ? pmlT_pre(gxmes, px[k].m0, px[k].q, px[k].a) : gxmes;
                    gzbv_o[k - 1] = gzbv;
                    gzmes_o[k - 1] = gzmes;
                    gxbu_o[k - 1] = gxbu;
                    gxmes_o[k - 1] = gxmes;
                }
            }
            szBV[zi][lx] = bv;
            szMes[zi][lx] = mz;
            if (k >= 1 && k < TZ / BLOCK_Y + 1) {
                sxBU[zi - 4][lx + 4] = bu;
                sxMes[zi - 4][lx + 4] = mx;
            }
        }
#pragma unroll
        for (int k = 0; k < KX; ++k) {
            if (!halo_round(k, 8 * TZ)) continue;
            const int e = tid + 256 * k;
            {
                // lanes past the halo store into the spare row TZ
                const bool valid = (8 * TZ) % 256 == 0 || e < 8 * TZ;
                const int zi = valid ? e >> 3 : TZ, h = e & 7, xi = h < 4 ? h : h + TILE_X;
                const int r = r0 + zi, cc = c0 - 4 + xi;
                T bu = 0, mx = 0;
                if (hok[k]) {
                    const bool inR = CZ(r <= nz - 2) && CX(cc >= 1 && cc <= nx - 2);
                    const T b = hb[k];
                    const T gxbu = inR ? g.dtdx * b * hu[k] : T(0);
                    const T gxmes = inR ? g.dtdx * b * hv[k] : T(0);
                    bu = ihs[k] >= 0 ? pmlT_pre(gxbu, hp[k].m1, hp[k].q, hp[k].a) : gxbu;
                    mx = ihs[k] >= 0 ? pmlT_pre(gxmes, hp[k].m0, hp[k].q, hp[k].a) : gxmes;
                }
                sxBU[zi][xi] = bu;
                sxMes[zi][xi] = mx;
            }
        }
    }
    };
    STFWI_DISPATCH(inx, inz, stage);
    // history frame and accumulators, issued after the staging (their
    // registers are not live across it) and before the barrier, whose
    // wait hides their latency
#pragma unroll
    for (int rr = 0; rr < RPT; ++rr) {
        const int r = r0 + ty + BLOCK_Y * rr;
        // history frame: read wherever the imaging below uses it (physical
        // cells; frame -1 is never used); other lanes read entry 0
        h0[rr] = h1[rr] = h2v[rr] = ac0[rr] = ac1[rr] = T(0);
        if (a.imaging) {
            const int rp = r - g.z0, cp = c - g.pad;
            const bool pok = rp >= 0 && rp < g.nzd && cp >= 0 && cp < g.nxd;
            const bool hok = hframe >= 0 && pok;
            const T *hfr = a.hist + (hh + (hframe >= a.hist_t0 ? hframe : a.hist_t0)) * 3 * g.hnp;
            const int o = hok ? rp * g.hpitch + cp + g.hcoff : 0, nph = (int)g.hnp;
            h0[rr] = ldgg(hfr, o);
            h1[rr] = ldgg(hfr, o + nph);
            h2v[rr] = ldgg(hfr, o + 2 * nph);
            // first two gradient accumulators of the cell (the history's layout)
            const T *accb = a.gacc + (i64)s * a.nacc * g.hnp;
            const int oa = pok ? rp * g.hpitch + cp + g.hcoff : 0;
            ac0[rr] = ldgg(accb, oa);
            if (a.gparams >= 2) ac1[rr] = ldgg(accb, oa + nph);
        }
    }
    __syncthreads();
    auto body = [&](auto tagx, auto tagz) {
        constexpr bool IX = decltype(tagx)::value, IZ = decltype(tagz)::value;
        // the column (x) term masks as per-lane weights, set once per tile
        // (a masked term adds 0 * a finite staged value)
        T wax1[NK], wax2[NK], wzi[NK], wmx1[NK], wmx2[NK];
#pragma unroll
        for (int k = 1; k <= NK; ++k) {
            const T ck = (T)fdc<FD>(k);
            const int i1 = c - k + 1, i2 = c + k;
            wax1[k - 1] = CX(c - k >= k - 1 && c - k <= nx - 1 - k) ? ck : T(0);
            wax2[k - 1] = CX(c + k - 1 >= k - 1 && c + k - 1 <= nx - 1 - k) ? -ck : T(0);
            wzi[k - 1] = CX(c >= 1 && c <= nx - 2) ? ck : T(0);
            wmx1[k - 1] = CX(k == 1 ? (i1 >= 1 && i1 <= nx - 2) : (i1 >= k && i1 <= nx - k)) ? ck : T(0);
            wmx2[k - 1] = CX(k == 1 ? (i2 >= 1 && i2 <= nx - 2) : (i2 >= k && i2 <= nx - k)) ? -ck : T(0);
        }
#pragma unroll
        for (int rr = 0; rr < RPT; ++rr) {
            const int lz = ty + BLOCK_Y * rr, r = r0 + lz;
            if ((!IZ && r >= nz) || (!IX && c >= nx)) continue;
            const int xc = lx + 4, zc = lz + 4;
            const int q = r * ld + c;
            T Ax = 0, Az = 0, Mes = 0;
            // the row (z) term masks as selects of the weights: no branches
            const bool rz = CZ(r <= nz - 2);
#pragma unroll
            for (int k = 1; k <= NK; ++k) {
                // (D+x)^T of Dx_(B U)_a (= L2M_ex_a = L_ez_a), rows [0, nz-2]
                Ax += (rz ? wax1[k - 1] : T(0)) * sxBU[lz][xc - k];
                Ax += (rz ? wax2[k - 1] : T(0)) * sxBU[lz][xc + k - 1];
                // (D+z)^T of Dz_(B V)_a (= L2M_ez_a = L_ex_a), columns [1, nx-2]
                const T z1 = CZ(r - k >= k - 1 && r - k <= nz - 1 - k) ? wzi[k - 1] : T(0);
                const T z2 = CZ(r + k - 1 >= k - 1 && r + k - 1 <= nz - 1 - k) ? -wzi[k - 1] : T(0);
                Az += z1 * szBV[zc - k][lx];
                Az += z2 * szBV[zc + k - 1][lx];
            }
#pragma unroll
            for (int k = 1; k <= NK; ++k) {
                // (D-x)^T of Dx_M_es_a, rows [0, nz-2]; term 1 cols [1, nx-2], k: [k, nx-k]
                Mes += (rz ? wmx1[k - 1] : T(0)) * sxMes[lz][xc - k + 1];
                Mes += (rz ? wmx2[k - 1] : T(0)) * sxMes[lz][xc + k];
            }
#pragma unroll
            for (int k = 1; k <= NK; ++k) {
                // (D-z)^T of Dz_M_es_a, columns [1, nx-2]; term 1 rows [1, nz-2], k: [k, nz-k]
                const int i1 = r - k + 1, i2 = r + k;
                const T m1 = CZ(k == 1 ? (i1 >= 1 && i1 <= nz - 2) : (i1 >= k && i1 <= nz - k)) ? wzi[k - 1] : T(0);
                const T m2 = CZ(k == 1 ? (i2 >= 1 && i2 <= nz - 2) : (i2 >= k && i2 <= nz - k)) ? -wzi[k - 1] : T(0);
                Mes += m1 * szMes[zc - k + 1][lx];
                Mes += m2 * szMes[zc + k][lx];
            }
            if (!IZ && g.fs && r == 0 && c >= 1 && c <= nx - 2) Mes += T(2 * fdc<FD>(1)) * szMes[zc][lx];

            const T l2m = l2m0[rr], mu = mu0[rr];
            const T lam = l2m - T(2) * mu;
            const T exa0 = exa0v[rr], eza0 = eza0v[rr];
            const T exan = exa0 + l2m * Ax + lam * Az;
            const T ezan = eza0 + lam * Ax + l2m * Az;
            const T esan = esa0v[rr] + T(2) * mu * Mes;
            stgg(exa, q, hasinj[rr] ? exan + inj2[rr] : exan);
            stgg(eza, q, hasinj[rr] ? ezan + inj3[rr] : ezan);
            stgg(esa, q, esan);

            // C-PML memories of this cell: psi' = b (G + psi), from the values
            // staged for row rr + 1 of the z-slab
            if (!IX) {
                const int ix = ixs[rr + 1];
                if (ix >= 0) {
                    T *pxn = a.psix_n + s * a.psix_stride + ix;
                    const PmlIn<T> &m = px[rr + 1];
                    pxn[2 * a.psix_len] = m.b * (gxmes_o[rr] + m.m0);
                    pxn[3 * a.psix_len] = m.b * (gxbu_o[rr] + m.m1);
                }
            }
            if (!IZ) {
                const int iz = izs[rr + 1];
                if (iz >= 0) {
                    T *pzn = a.psiz_n + s * a.psiz_stride + iz;
                    const PmlIn<T> &m = pz[rr + 1];
                    __builtin_nontemporal_store(m.b * (gzbv_o[rr] + m.m0), pzn + 2 * a.psiz_len);
                    __builtin_nontemporal_store(m.b * (gzmes_o[rr] + m.m1), pzn + 3 * a.psiz_len);
                }
            }

            if (a.snap || a.aen) {
                const T vals[5] = {Ua[q], Va[q], exan, ezan, esan};
                if (a.snap) {
                    int k = 0;
                    for (int f = 0; f < 5; ++f)
                        if (a.snap_mask & (1 << f)) {
                            stst(a.snap + (((i64)s * a.nsnap + k) * a.nt_max + it) * nz * nx, r * nx + c, vals[f]);
                            ++k;
                        }
                }
                if (a.aen) {
                    // adjoint energy of one component (the Wr of eprecond1)
                    const T v = vals[a.aen_comp];
                    T *e = a.aen + (i64)s * nz * nx + (i64)r * nx + c;
                    *e = *e + v * v;
                }
            }

            // imaging on the physical cells
            const int rp = r - g.z0, cp = c - g.pad;
            if (a.imaging == 0 || !(CZ(rp >= 0 && rp < g.nzd) && CX(cp >= 0 && cp < g.nxd))) continue;
            const int ip = rp * g.nxd + cp;
            T *acc = a.gacc + (i64)s * a.nacc * g.hnp + (rp * g.hpitch + cp + g.hcoff);
            if (a.imaging == 2) {
                // exact discrete gradient: model dependence of step it, paired with
                // its carry-in strains (history frame it-1)
                T gl2m = 0, gl = 0, gm = 0, gb = 0;
                if (it > 0) {
                    const T exm = h0[rr], ezm = h1[rr], esm = h2v[rr];
                    gl2m = fma(exm, Ax, ezm * Az);
                    gl = fma(ezm, Ax, exm * Az);
                    gm = T(2) * esm * Mes;
                }
                if (!IZ && g.fs && r == 0 && c <= nx - 2) {
                    // sigma_zz = 0 row: Dz_V = L_ratio Dx_U dz/dx; row 0 has no z-PML
                    const T tt = g.dtdz * eza0 * a.dxu0[(hs + it) * nx + c] * g.dz / g.dx;
                    gl += -tt / l2m;
                    gl2m += tt * lam / (l2m * l2m);
                }
                if (sh.type == 1 && r == sh.sz && c == sh.sx) {
                    const T lm = lam + mu;
                    const T stm = a.wav[hs + it] * g.dt / (T(2) * lm * g.dx * g.dz);
                    const T Qm = -stm / lm * (exa0 + eza0);
                    gl += Qm;
                    gm += Qm;
                }
                if (a.gparams == 3) {
                    const T *h2 = a.hist2 + (hh + it) * 2 * g.np + ip;
                    gb = fma(Ua[q], h2[0], Va[q] * h2[g.np]);
                    acc[0] = ac0[rr] + gl2m;
                    acc[g.hnp] = ac1[rr] + gl;
                    acc[2 * g.hnp] += gm;
                    acc[3 * g.hnp] += gb;
                } else {
                    // a wavefront's 64 cells fill two whole 128-byte lines of the
                    // accumulator rows: stored without keeping them in L2
                    __builtin_nontemporal_store(ac0[rr] + (gm - T(2) * gl), acc);
                    if (a.gparams == 2) __builtin_nontemporal_store(ac1[rr] + (gl2m + gl), acc + g.hnp);
                }
            } else {
                // Zhou imaging (fwi/imaging.py:125-152): forward frame it with the
                // adjoint carry-out of step it; centred periodic differences of B v_a
                const T ex = h0[rr], ez = h1[rr], es = h2v[rr];
                const T *Bi = a.Bimg + g.origin;
                const int cpl = c + 1 == nx ? 0 : c + 1, cmi = c == 0 ? nx - 1 : c - 1;
                const int rpl = r + 1 == nz ? 0 : r + 1, rmi = r == 0 ? nz - 1 : r - 1;
                const int qxp = r * ld + cpl, qxm = r * ld + cmi;
                const int qzp = rpl * ld + c, qzm = rmi * ld + c;
                const T i2dx = T(2) * g.dx, i2dz = T(2) * g.dz;
                const T dxBU = fma(Bi[qxp], Ua[qxp], -(Bi[qxm] * Ua[qxm])) / i2dx;
                const T dzBV = fma(Bi[qzp], Va[qzp], -(Bi[qzm] * Va[qzm])) / i2dz;
                const T dzBU = fma(Bi[qzp], Ua[qzp], -(Bi[qzm] * Ua[qzm])) / i2dz;
                const T dxBV = fma(Bi[qxp], Va[qxp], -(Bi[qxm] * Va[qxm])) / i2dx;
                const T glam = (ex + ez) * (dxBU + dzBV);
                const T gmu = fma(es, dzBU + dxBV, fma(T(2) * ez, dzBV, T(2) * ex * dxBU));
                if (a.gparams == 3) {
                    // density term: U_a(t) pairs with v(t-1) - v(t+1) (summation by
                    // parts of sum_{t=1}^{nt-2} v(t) (U_a(t+1) - U_a(t-1)) / 2dt)
                    T grl = 0;
                    const T i2dt = T(2) * g.dt;
                    if (it >= 2 && it <= sh.nt - 1) {
                        const T *h2 = a.hist2 + (hh + it - 1) * 2 * g.np + ip;
                        grl += Ua[q] * h2[0] / i2dt + Va[q] * h2[g.np] / i2dt;
                    }
                    if (it <= sh.nt - 3) {
                        const T *h2 = a.hist2 + (hh + it + 1) * 2 * g.np + ip;
                        grl -= Ua[q] * h2[0] / i2dt + Va[q] * h2[g.np] / i2dt;
                    }
                    acc[0] = ac0[rr] + glam;
                    acc[g.hnp] = ac1[rr] + gmu;
                    acc[2 * g.hnp] += grl;
                } else {
                    __builtin_nontemporal_store(ac0[rr] + (gmu - T(2) * glam), acc);
                    if (a.gparams == 2) __builtin_nontemporal_store(ac1[rr] + glam, acc + g.hnp);
                }
            }
        }
    };
    STFWI_DISPATCH(inx, inz, body);
}


// history frame `frame` of every shot from the strains held in the state
// buffers (the state restored from a checkpoint is the carry-in of the
// segment's first step)
template <typename T>
__global__ void k_hist_from_state(Geo<T> g, const T *state, i64 sstride, T *hist, const ShotDev *shots, int frame,
                                  int shot0)
{
    const int s = shot0 + blockIdx.y;
    const int hoff = shots[s].hoff;
    const i64 np = g.np;
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < np; i += (i64)gridDim.x * blockDim.x) {
        const int rp = (int)(i / g.nxd), cp = (int)(i - (i64)rp * g.nxd);
        const i64 cell = g.origin + (i64)(rp + g.z0) * g.ld + cp + g.pad;
        const T *st = state + s * sstride;
        T *h = hist + ((i64)hoff + frame) * 3 * g.hnp + (i64)rp * g.hpitch + cp + g.hcoff;
        h[0] = st[2 * g.fstride + cell];
        h[g.hnp] = st[3 * g.fstride + cell];
        h[2 * g.hnp] = st[4 * g.fstride + cell];
    }
}

// sum of per-shot accumulators and chain rule to (Vs, Vp, rho); physical
// cell i sits at (i / nxd) apitch + i % nxd + acoff of each accumulator
// plane of anp elements
template <typename T, typename C>
__global__ void k_finalize(int n_shots, i64 np, int nacc, int imaging, int gparams, T dt,
                           const T *__restrict__ gacc, const C *__restrict__ vs, const C *__restrict__ vp,
                           const C *__restrict__ rho, T *out, int nxd, int apitch, int acoff, i64 anp)
{
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const i64 ia = (i / nxd) * apitch + (i % nxd) + acoff;
    T s[4] = {0, 0, 0, 0};
    for (int sh = 0; sh < n_shots; ++sh)
        for (int k = 0; k < nacc; ++k) s[k] += gacc[((i64)sh * nacc + k) * anp + ia];
    const C v_s = vs[i], v_p = vp[i], d = rho[i];
    T gvs = 0, gvp = 0, grho = 0;
    if (imaging == 1) {
        for (int k = 0; k < 4; ++k) s[k] = -dt * s[k];
        if (gparams == 3) {
            gvs = T(C(2) * d * v_s) * (s[1] - T(2) * s[0]);
            gvp = T(C(2) * d * v_p) * s[0];
            grho = T(v_p * v_p - C(2) * v_s * v_s) * s[0] + T(v_s * v_s) * s[1] + s[2];
        } else {
            gvs = T(C(2) * d * v_s) * s[0];
            if (gparams == 2) gvp = T(C(2) * d * v_p) * s[1];
        }
    } else {
        if (gparams == 3) {
            const T tvs = T(v_s), tvp = T(v_p), td = T(d);
            gvs = T(2) * td * tvs * (s[2] - T(2) * s[1]);
            gvp = T(2) * td * tvp * (s[0] + s[1]);
            grho = tvs * tvs * s[2] + tvp * tvp * s[0] + (tvp * tvp - T(2) * tvs * tvs) * s[1] - s[3] / (td * td);
        } else {
            gvs = T(2) * T(d) * T(v_s) * s[0];
            if (gparams == 2) gvp = T(2) * T(d) * T(v_p) * s[1];
        }
    }
    out[i] = gvs;
    out[np + i] = gvp;
    out[2 * np + i] = grho;
}

// ---------------------------------------------------------------------------
// residual prep: taper, DF-I biquad cascade, misfit and its transpose
// (fwi/gradient.py:139-156, fwi/filters.py:14-42); one block per shot
// ---------------------------------------------------------------------------
#define MAX_SOS 8
template <typename T>
struct ResArgs {
    const ShotDev *shots;
    const T *pred;        // [shot][4][nt_max][nrec_max]
    const T *obs;         // [shot][nt_max][nrec_max]
    T *rec_adj;           // [shot][4][nt_max][nrec_max]
    T *loss;              // [shot]
    int nt_max, nrec_max, comp, n_sos, n_taper;
    T coef[MAX_SOS][5];   // b0 b1 b2 a1 a2 (normalised by a0)
    T w;                  // scale / n_shots
};

template <typename T>
__device__ __forceinline__ T taper_w(int t, int nt, int n_taper)
{
    const int n = n_taper < nt ? n_taper : nt;
    if (n <= 0 || t < nt - n) return T(1);
    const int k = t - (nt - n) + 1;   // 1..n
    return (T)((cos((double)k / n * M_PI) + 1.0) / 2.0);
}

// one thread per (shot, receiver): the component channel's residual, its
// transpose in place, and the receiver's sum of squares in lpart
// [shot][nrec_max] (0 for absent receivers)
template <typename T>
__global__ __launch_bounds__(64) void k_residual(ResArgs<T> a, double *lpart)
{
    const int s = blockIdx.y;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.nrec_max) return;
    const ShotDev sh = a.shots[s];
    const int nt = sh.nt;
    const i64 chs = (i64)a.nt_max * a.nrec_max;
    const T *__restrict__ pred = a.pred + ((i64)s * 4 + a.comp) * chs;
    const T *__restrict__ obs = a.obs + (i64)s * chs;
    T *__restrict__ rr = a.rec_adj + ((i64)s * 4 + a.comp) * chs + j;
    if (j >= sh.nrec) {
        for (int t = 0; t < a.nt_max; ++t) rr[(i64)t * a.nrec_max] = 0;
        lpart[(i64)s * a.nrec_max + j] = 0;
        return;
    }
    // forward: r = F T pred - F T obs, stored in place
    T xs[MAX_SOS][2][2], ys[MAX_SOS][2][2];   // [section][pred/obs][lag]
    for (int k = 0; k < a.n_sos; ++k)
        for (int u = 0; u < 2; ++u) xs[k][u][0] = xs[k][u][1] = ys[k][u][0] = ys[k][u][1] = 0;
    T ssq = 0;
#pragma unroll 4
    for (int t = 0; t < nt; ++t) {
        const T tw = taper_w<T>(t, nt, a.n_taper);
        T v[2] = {pred[(i64)t * a.nrec_max + j], obs[(i64)t * a.nrec_max + j]};
        for (int u = 0; u < 2; ++u) {
            if (a.n_taper > 0) v[u] = v[u] * tw;
            for (int k = 0; k < a.n_sos; ++k) {
                const T *cf = a.coef[k];
                const T y = cf[0] * v[u] + cf[1] * xs[k][u][0] + cf[2] * xs[k][u][1] - cf[3] * ys[k][u][0] - cf[4] * ys[k][u][1];
                xs[k][u][1] = xs[k][u][0];
                xs[k][u][0] = v[u];
                ys[k][u][1] = ys[k][u][0];
                ys[k][u][0] = y;
                v[u] = y;
            }
        }
        const T res = v[0] - v[1];
        rr[(i64)t * a.nrec_max] = res;
        ssq += res * res;
    }
    for (int t = nt; t < a.nt_max; ++t) rr[(i64)t * a.nrec_max] = 0;
    lpart[(i64)s * a.nrec_max + j] = (double)ssq;
    // transpose: cot = 2 r w / N, then F^T (reverse-time recursion), T^T
    const T g0 = T(2) * (a.w / (T)((double)nt * sh.nrec));
    T cx1[MAX_SOS], cx2[MAX_SOS], cy1[MAX_SOS], cy2[MAX_SOS];
    for (int k = 0; k < a.n_sos; ++k) cx1[k] = cx2[k] = cy1[k] = cy2[k] = 0;
    for (int t = nt - 1; t >= 0; --t) {
        T gcur = rr[(i64)t * a.nrec_max] * g0;
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
        if (a.n_taper > 0) gcur = gcur * taper_w<T>(t, nt, a.n_taper);
        rr[(i64)t * a.nrec_max] = gcur;
    }
}

// Without a filter the residual and its transpose are elementwise in time:
// r = T pred - T obs, then 2 r w / N and T^T, the values k_residual stores.
// Thread (receiver j, block ch of RES_TCH steps) writes its steps and its
// block's sum of squares into lpart [shot][chunk][nrec_max].
#define RES_TCH 64
template <typename T>
__global__ __launch_bounds__(64) void k_residual_el(ResArgs<T> a, double *lpart)
{
    const int s = blockIdx.z, ch = blockIdx.y, nch = gridDim.y;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.nrec_max) return;
    const ShotDev sh = a.shots[s];
    const int nt = sh.nt;
    const i64 chs = (i64)a.nt_max * a.nrec_max;
    const T *__restrict__ pred = a.pred + ((i64)s * 4 + a.comp) * chs + j;
    const T *__restrict__ obs = a.obs + (i64)s * chs + j;
    T *__restrict__ rr = a.rec_adj + ((i64)s * 4 + a.comp) * chs + j;
    const int t0 = ch * RES_TCH, t1 = t0 + RES_TCH < a.nt_max ? t0 + RES_TCH : a.nt_max;
    double *lp = lpart + ((i64)s * nch + ch) * a.nrec_max + j;
    if (j >= sh.nrec) {
        for (int t = t0; t < t1; ++t) rr[(i64)t * a.nrec_max] = 0;
        *lp = 0;
        return;
    }
    const T g0 = T(2) * (a.w / (T)((double)nt * sh.nrec));
    T ssq = 0;
#pragma unroll 4
    for (int t = t0; t < t1; ++t) {
        T out = 0;
        if (t < nt) {
            T v0 = pred[(i64)t * a.nrec_max], v1 = obs[(i64)t * a.nrec_max];
            const T tw = taper_w<T>(t, nt, a.n_taper);
            if (a.n_taper > 0) {
                v0 = v0 * tw;
                v1 = v1 * tw;
            }
            const T res = v0 - v1;
            ssq += res * res;
            out = res * g0;
            if (a.n_taper > 0) out = out * tw;
        }
        rr[(i64)t * a.nrec_max] = out;
    }
    *lp = (double)ssq;
}

// the other three channels of the receiver adjoints are zero
template <typename T>
__global__ void k_residual_zero(T *rec_adj, i64 n, i64 chs, int comp)
{
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        if ((i / chs) % 4 != comp) rec_adj[i] = 0;
}

// per-shot misfit from lpart [shot][nch][nrec_max]: thread t sums receivers
// t, t + 256, ... in order (each over its nch time blocks in order), then a
// fixed tree (one block per shot)
template <typename T>
__global__ __launch_bounds__(256) void k_residual_loss(ResArgs<T> a, const double *lpart, int nch)
{
    const int s = blockIdx.x;
    const ShotDev sh = a.shots[s];
    __shared__ double red[256];
    double part = 0;
    for (int j = threadIdx.x; j < a.nrec_max; j += blockDim.x)
        for (int c = 0; c < nch; ++c) part += lpart[((i64)s * nch + c) * a.nrec_max + j];
    red[threadIdx.x] = part;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.loss[s] = (T)((double)a.w * (red[0] / ((double)sh.nt * sh.nrec)));
}

}  // namespace stfwi
