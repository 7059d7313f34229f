/*
 * CPU restatement of the reference's elastic velocity-strain forward kernel,
 * its hand-coded discrete adjoint, and the two imaging conditions.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle that the HIP path is checked
 * against; nothing in the product (stfwi_amd/) may include or link it.
 *
 * The file is a template: orc.c includes it once per element type with
 *   T      element type (float, double, double _Complex)
 *   SFX    symbol suffix (f32, f64, c128)
 *   RE(x)  real part of a T
 * so the same restatement gives the FP32 CPU baseline, the FP64 oracle and a
 * complex128 build for complex-step differentiation.
 *
 * Every array operation follows the reference slice by slice, in the same
 * order, so that FP64 results agree with the reference to rounding:
 *   forward step        fwi/forward.py:250-469
 *   model embedding     fwi/forward.py:193-238
 *   C-PML coefficients  fwi/forward.py:35-112
 *   adjoint step        fwi/adjoint_jax_kernel.py:189-476
 *   Zhou imaging        fwi/imaging.py:83-169
 *   exact gradient      the discrete transpose of the forward's model
 *                       dependence (what jax.grad computes for
 *                       fwi/gradient.py:106-132 through fwi/loss.py:33-152)
 */

#define CAT_(a, b) a##_##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SFX)

/* ------------------------------------------------------------------ */
/* whole-array derivative with per-term regions (reference slicing)     */
/* axis 0 = z (rows), 1 = x (columns).  plus: D+ (f[i+k]-f[i-k+1]),     */
/* minus: D- (f[i+k-1]-f[i-k]).  Term 1 covers [lo1,hi1] along the axis; */
/* term k>=2 covers the in-bounds range ([k-1,n-1-k] for D+, [k,n-k]    */
/* for D-).  The other axis covers [olo,ohi].  out is zeroed first.     */
/* ------------------------------------------------------------------ */
static void FN(deriv)(T *out, const T *f, int nz, int nx, int axis, int plus,
                      int lo1, int hi1, int olo, int ohi,
                      const double *c, int nk)
{
    const long N = (long)nz * nx;
    for (long i = 0; i < N; ++i) out[i] = 0;
    const int n = axis == 0 ? nz : nx;
    for (int k = 1; k <= nk; ++k) {
        const int lo = (k == 1) ? lo1 : (plus ? k - 1 : k);
        const int hi = (k == 1) ? hi1 : (plus ? n - 1 - k : n - k);
        const int a = plus ? k : k - 1;
        const int b = plus ? -(k - 1) : -k;
        const T ck = (T)c[k];
        if (axis == 0) {
#pragma omp parallel for schedule(static)
            for (int i = lo; i <= hi; ++i)
                for (int o = olo; o <= ohi; ++o)
                    out[(long)i * nx + o] += ck * (f[(long)(i + a) * nx + o] - f[(long)(i + b) * nx + o]);
        } else {
#pragma omp parallel for schedule(static)
            for (int o = olo; o <= ohi; ++o)
                for (int i = lo; i <= hi; ++i)
                    out[(long)o * nx + i] += ck * (f[(long)o * nx + i + a] - f[(long)o * nx + i + b]);
        }
    }
}

/* transpose of deriv(): accumulate into out (not zeroed).  For each term
 * the "+" scatter over the whole region is done before the "-" scatter,
 * as in fwi/adjoint_jax_kernel.py (two .at[].add per term). */
static void FN(deriv_T)(T *out, const T *Da, int nz, int nx, int axis, int plus,
                        int lo1, int hi1, int olo, int ohi,
                        const double *c, int nk)
{
    const int n = axis == 0 ? nz : nx;
    for (int k = 1; k <= nk; ++k) {
        const int lo = (k == 1) ? lo1 : (plus ? k - 1 : k);
        const int hi = (k == 1) ? hi1 : (plus ? n - 1 - k : n - k);
        const int a = plus ? k : k - 1;
        const int b = plus ? -(k - 1) : -k;
        const T cp = (T)c[k], cm = (T)(-c[k]);
        for (int pass = 0; pass < 2; ++pass) {
            const int sh = pass == 0 ? a : b;
            const T ck = pass == 0 ? cp : cm;
            if (axis == 0) {
                /* parallel over target rows t = i + sh */
#pragma omp parallel for schedule(static)
                for (int t = lo + sh; t <= hi + sh; ++t)
                    for (int o = olo; o <= ohi; ++o)
                        out[(long)t * nx + o] += ck * Da[(long)(t - sh) * nx + o];
            } else {
#pragma omp parallel for schedule(static)
                for (int o = olo; o <= ohi; ++o)
                    for (int i = lo; i <= hi; ++i)
                        out[(long)o * nx + i + sh] += ck * Da[(long)o * nx + i];
            }
        }
    }
}

/* C-PML recursion on the whole grid (fwi/forward.py:107-112).  Profiles are
 * 1-D: per column for axis x, per row for axis z. */
static void FN(pml)(T *D, T *psi, int nz, int nx, int axis,
                    const double *k, const double *b, const double *a)
{
#pragma omp parallel for schedule(static)
    for (int r = 0; r < nz; ++r)
        for (int cidx = 0; cidx < nx; ++cidx) {
            const int j = axis == 0 ? r : cidx;
            const long p = (long)r * nx + cidx;
            const T pn = (T)b[j] * psi[p] + (T)a[j] * D[p];
            psi[p] = pn;
            D[p] = D[p] / (T)k[j] + pn;
        }
}

/* transpose of the C-PML recursion (fwi/adjoint_jax_kernel.py:51-54) */
static void FN(pml_T)(T *Da, T *psia, int nz, int nx, int axis,
                      const double *k, const double *b, const double *a)
{
#pragma omp parallel for schedule(static)
    for (int r = 0; r < nz; ++r)
        for (int cidx = 0; cidx < nx; ++cidx) {
            const int j = axis == 0 ? r : cidx;
            const long p = (long)r * nx + cidx;
            const T dn = Da[p], pa = psia[p];
            Da[p] = ((T)(1.0 / k[j]) + (T)a[j]) * dn + (T)a[j] * pa;
            psia[p] = (T)b[j] * (dn + pa);
        }
}

/* ------------------------------------------------------------------ */
/* model set-up                                                         */
/* ------------------------------------------------------------------ */
typedef struct {
    int nz, nx, z0;            /* padded grid, first physical row */
    T *L2M, *L, *M, *B, *Lr;   /* full padded arrays; Lr only with FS */
    double *kx, *bx, *ax, *kz, *bz, *az;
    double c[5];
    int nk;
} FN(orc_model);

/* nearest-neighbour embedding of a physical field into the padded grid
 * (fwi/forward.py:195-212) */
static void FN(embed)(T *full, const T *phys, const orc_geom *g, int nz, int nx, int z0)
{
    const int pad = g->pad, nzd = g->nz_dom, nxd = g->nx_dom;
    for (long i = 0; i < (long)nz * nx; ++i) full[i] = 0;
    for (int i = 0; i < nzd; ++i)
        for (int j = 0; j < nxd; ++j) full[(long)(z0 + i) * nx + pad + j] = phys[(long)i * nxd + j];
    /* the copies into the PML carry no derivative (stop_gradient,
     * fwi/forward.py:214-219): for the complex-step build they keep only the
     * real part; for real types RE() is the identity */
    for (int i = 0; i < nzd; ++i) {
        T *row = full + (long)(z0 + i) * nx;
        for (int j = 0; j < pad; ++j) row[j] = (T)RE(row[pad]);
        for (int j = nx - pad; j < nx; ++j) row[j] = (T)RE(row[pad + nxd - 1]);
    }
    for (int i = z0 + nzd; i < nz; ++i)
        for (int j = 0; j < nx; ++j) full[(long)i * nx + j] = (T)RE(full[(long)(z0 + nzd - 1) * nx + j]);
    if (!g->free_surface)
        for (int i = 0; i < pad; ++i)
            for (int j = 0; j < nx; ++j) full[(long)i * nx + j] = (T)RE(full[(long)pad * nx + j]);
}

static int FN(model_init)(FN(orc_model) *m, const orc_geom *g, const T *vs, const T *vp, const T *rho)
{
    const int pad = g->pad;
    m->nz = g->nz_dom + (g->free_surface ? pad : 2 * pad);
    m->nx = g->nx_dom + 2 * pad;
    m->z0 = g->free_surface ? 0 : pad;
    const int nz = m->nz, nx = m->nx;
    const long N = (long)nz * nx;
    T *vsf = malloc(N * sizeof(T)), *vpf = malloc(N * sizeof(T)), *rhof = malloc(N * sizeof(T));
    m->L2M = malloc(N * sizeof(T)); m->L = malloc(N * sizeof(T)); m->M = malloc(N * sizeof(T));
    m->B = malloc(N * sizeof(T)); m->Lr = malloc(N * sizeof(T));
    m->kx = malloc(nx * sizeof(double)); m->bx = malloc(nx * sizeof(double)); m->ax = malloc(nx * sizeof(double));
    m->kz = malloc(nz * sizeof(double)); m->bz = malloc(nz * sizeof(double)); m->az = malloc(nz * sizeof(double));
    FN(embed)(vsf, vs, g, nz, nx, m->z0);
    FN(embed)(vpf, vp, g, nz, nx, m->z0);
    FN(embed)(rhof, rho, g, nz, nx, m->z0);
    for (long i = 0; i < N; ++i) {
        m->M[i] = rhof[i] * (vsf[i] * vsf[i]);
        m->L2M[i] = rhof[i] * (vpf[i] * vpf[i]);
        m->L[i] = m->L2M[i] - (T)2 * m->M[i];
        m->B[i] = (T)1 / rhof[i];
        m->Lr[i] = -m->L[i] / m->L2M[i];
    }
    free(vsf); free(vpf); free(rhof);

    /* FD coefficients (fwi/forward.py:177-182) */
    m->c[0] = 0;
    if (g->fd_order == 8) {
        m->c[1] = 1225.0 / 1024.0; m->c[2] = -245.0 / 3072.0; m->c[3] = 49.0 / 5120.0; m->c[4] = -5.0 / 7168.0; m->nk = 4;
    } else if (g->fd_order >= 4) {
        m->c[1] = 9.0 / 8.0; m->c[2] = -1.0 / 24.0; m->c[3] = 0; m->c[4] = 0; m->nk = 2;
    } else {
        m->c[1] = 1.0; m->c[2] = m->c[3] = m->c[4] = 0; m->nk = 1;
    }

    /* C-PML profiles (fwi/forward.py:35-104); Vp_max over the physical Vp */
    double vpmax = -1e300;
    for (long i = 0; i < (long)g->nz_dom * g->nx_dom; ++i)
        if (RE(vp[i]) > vpmax) vpmax = RE(vp[i]);
    const double L_pml = pad * g->dx;
    const double d0 = -(2 + 1) * vpmax * log(1e-3) / (2 * L_pml);
    double *dpx = calloc(nx, sizeof(double)), *kpx = malloc(nx * sizeof(double)), *apx = calloc(nx, sizeof(double));
    double *dpz = calloc(nz, sizeof(double)), *kpz = malloc(nz * sizeof(double)), *apz = calloc(nz, sizeof(double));
    for (int j = 0; j < nx; ++j) kpx[j] = 1.0;
    for (int j = 0; j < nz; ++j) kpz[j] = 1.0;
    for (int i = 0; i < pad; ++i) {
        const double dist = (i + 1) * g->dx / L_pml;
        const double dd = d0 * (dist * dist), kk = 1.0 + (2.0 - 1.0) * (dist * dist);
        const double al = M_PI * g->fc * (1.0 - dist);
        int col = pad - 1 - i;
        dpx[col] = dd; kpx[col] = kk; apx[col] = al;
        col = nx - pad + i;
        dpx[col] = dd; kpx[col] = kk; apx[col] = al;
        const int bottom_start = g->free_surface ? g->nz_dom : pad + g->nz_dom;
        int row = bottom_start + i;
        dpz[row] = dd; kpz[row] = kk; apz[row] = al;
        if (!g->free_surface) {
            row = pad - 1 - i;
            dpz[row] = dd; kpz[row] = kk; apz[row] = al;
        }
    }
    for (int j = 0; j < nx; ++j) {
        m->kx[j] = kpx[j];
        m->bx[j] = exp(-(dpx[j] / kpx[j] + apx[j]) * g->dt);
        const double den = kpx[j] * (dpx[j] + kpx[j] * apx[j]);
        m->ax[j] = den > 1e-10 ? dpx[j] * (m->bx[j] - 1.0) / den : 0.0;
    }
    for (int j = 0; j < nz; ++j) {
        m->kz[j] = kpz[j];
        m->bz[j] = exp(-(dpz[j] / kpz[j] + apz[j]) * g->dt);
        const double den = kpz[j] * (dpz[j] + kpz[j] * apz[j]);
        m->az[j] = den > 1e-10 ? dpz[j] * (m->bz[j] - 1.0) / den : 0.0;
    }
    free(dpx); free(kpx); free(apx); free(dpz); free(kpz); free(apz);
    return 0;
}

static void FN(model_free)(FN(orc_model) *m)
{
    free(m->L2M); free(m->L); free(m->M); free(m->B); free(m->Lr);
    free(m->kx); free(m->bx); free(m->ax); free(m->kz); free(m->bz); free(m->az);
}

/* export the padded moduli and PML profiles (used by tests to pin the
 * set-up against closed forms) */
int FN(orc_setup)(const orc_geom *g, const T *vs, const T *vp, const T *rho,
                  T *L2M, T *L, T *M, T *B, double *prof /* 3*nx + 3*nz: kx bx ax kz bz az */)
{
    FN(orc_model) m;
    FN(model_init)(&m, g, vs, vp, rho);
    const long N = (long)m.nz * m.nx;
    if (L2M) memcpy(L2M, m.L2M, N * sizeof(T));
    if (L) memcpy(L, m.L, N * sizeof(T));
    if (M) memcpy(M, m.M, N * sizeof(T));
    if (B) memcpy(B, m.B, N * sizeof(T));
    if (prof) {
        memcpy(prof, m.kx, m.nx * sizeof(double));
        memcpy(prof + m.nx, m.bx, m.nx * sizeof(double));
        memcpy(prof + 2 * m.nx, m.ax, m.nx * sizeof(double));
        memcpy(prof + 3 * m.nx, m.kz, m.nz * sizeof(double));
        memcpy(prof + 3 * m.nx + m.nz, m.bz, m.nz * sizeof(double));
        memcpy(prof + 3 * m.nx + 2 * m.nz, m.az, m.nz * sizeof(double));
    }
    FN(model_free)(&m);
    return 0;
}

/* ------------------------------------------------------------------ */
/* forward (fwi/forward.py:119-501)                                     */
/* rec:     4 x nt x nrec (vx, vz, exx, ezz) or NULL                    */
/* h_uv:    nt x 2 x nz x nx   U, V after each step           or NULL   */
/* h_e:     nt x 3 x nz x nx   exx, ezz, exz after each step  or NULL   */
/* h_s:     nt x 2 x nz x nx   dU/dB, dV/dB of each step      or NULL   */
/* h_dxu0:  nt x nx            C-PML'd D-x U on row 0         or NULL   */
/* ------------------------------------------------------------------ */
int FN(orc_forward)(const orc_geom *g, const T *vs, const T *vp, const T *rho,
                    int sx_dom, int sz_dom, int src_type, const T *wav,
                    int nrec, const int *rx_dom, const int *rz_dom,
                    T *rec, T *h_uv, T *h_e, T *h_s, T *h_dxu0)
{
    FN(orc_model) m;
    FN(model_init)(&m, g, vs, vp, rho);
    const int nz = m.nz, nx = m.nx, nt = g->nt, fs = g->free_surface, nk = m.nk;
    const long N = (long)nz * nx;
    const double *c = m.c;
    const int sx = sx_dom + g->pad, sz = sz_dom + m.z0;
    T *U = calloc(N, sizeof(T)), *V = calloc(N, sizeof(T));
    T *ex = calloc(N, sizeof(T)), *ez = calloc(N, sizeof(T)), *es = calloc(N, sizeof(T));
    T *p[10];
    for (int i = 0; i < 10; ++i) p[i] = calloc(N, sizeof(T));
    T *f1 = malloc(N * sizeof(T)), *f2 = malloc(N * sizeof(T)), *f3 = malloc(N * sizeof(T));
    T *D1 = malloc(N * sizeof(T)), *D2 = malloc(N * sizeof(T)), *D3 = malloc(N * sizeof(T));
    const T dtdx = (T)(g->dt / g->dx), dtdz = (T)(g->dt / g->dz), half = (T)0.5;
    const T fsrc = (T)(g->dt / (g->dx * g->dz));

    for (int it = 0; it < nt; ++it) {
        /* 1) U */
        for (long i = 0; i < N; ++i) { f1[i] = m.L2M[i] * ex[i]; f2[i] = m.L[i] * ez[i]; f3[i] = (T)2 * m.M[i] * es[i]; }
        FN(deriv)(D1, f1, nz, nx, 1, 1, 0, nx - 2, 0, nz - 2, c, nk);
        FN(pml)(D1, p[0], nz, nx, 1, m.kx, m.bx, m.ax);
        FN(deriv)(D2, f2, nz, nx, 1, 1, 0, nx - 2, 0, nz - 2, c, nk);
        FN(pml)(D2, p[1], nz, nx, 1, m.kx, m.bx, m.ax);
        FN(deriv)(D3, f3, nz, nx, 0, 0, 1, nz - 2, 1, nx - 2, c, nk);
        if (fs)
            for (int j = 1; j <= nx - 2; ++j) D3[j] = (T)(2 * c[1]) * f3[j];
        FN(pml)(D3, p[2], nz, nx, 0, m.kz, m.bz, m.az);
        for (int r = 0; r <= nz - 2; ++r)
            for (int j = 1; j <= nx - 2; ++j) {
                const long q = (long)r * nx + j;
                if (h_s) h_s[((long)it * 2 + 0) * N + q] = dtdx * D1[q] + dtdx * D2[q] + dtdz * D3[q];
                U[q] = U[q] + dtdx * m.B[q] * D1[q] + dtdx * m.B[q] * D2[q] + dtdz * m.B[q] * D3[q];
            }
        /* 2) V */
        for (long i = 0; i < N; ++i) { f1[i] = m.L2M[i] * ez[i]; f2[i] = m.L[i] * ex[i]; }
        FN(deriv)(D1, f3, nz, nx, 1, 0, 1, nx - 2, 0, nz - 2, c, nk);            /* Dx_M_es */
        FN(pml)(D1, p[5], nz, nx, 1, m.kx, m.bx, m.ax);
        FN(deriv)(D2, f1, nz, nx, 0, 1, 0, nz - 2, 1, nx - 2, c, nk);            /* Dz_L2M_ez */
        FN(pml)(D2, p[3], nz, nx, 0, m.kz, m.bz, m.az);
        FN(deriv)(D3, f2, nz, nx, 0, 1, 0, nz - 2, 1, nx - 2, c, nk);            /* Dz_L_ex */
        FN(pml)(D3, p[4], nz, nx, 0, m.kz, m.bz, m.az);
        for (int r = 0; r <= nz - 2; ++r)
            for (int j = 1; j <= nx - 2; ++j) {
                const long q = (long)r * nx + j;
                if (h_s) h_s[((long)it * 2 + 1) * N + q] = dtdx * D1[q] + dtdz * D2[q] + dtdz * D3[q];
                V[q] = V[q] + dtdx * m.B[q] * D1[q] + dtdz * m.B[q] * D2[q] + dtdz * m.B[q] * D3[q];
            }
        const long s = (long)sz * nx + sx;
        if (src_type == 0) {
            V[s] += fsrc * m.B[s] * wav[it];
            if (h_s) h_s[((long)it * 2 + 1) * N + s] += fsrc * wav[it];
        }
        /* 3) exx */
        FN(deriv)(D1, U, nz, nx, 1, 0, 1, nx - 1, 0, nz - 1, c, nk);             /* Dx_U */
        FN(pml)(D1, p[6], nz, nx, 1, m.kx, m.bx, m.ax);
        for (int r = 0; r < nz; ++r)
            for (int j = 1; j < nx; ++j) ex[(long)r * nx + j] += dtdx * D1[(long)r * nx + j];
        if (h_dxu0) for (int j = 0; j < nx; ++j) h_dxu0[(long)it * nx + j] = D1[j];
        /* 4) ezz (+ free-surface row) */
        FN(deriv)(D2, V, nz, nx, 0, 0, 1, nz - 1, 0, nx - 2, c, nk);             /* Dz_V */
        if (fs)
            for (int j = 0; j <= nx - 2; ++j) D2[j] = m.Lr[j] * D1[j] * (T)g->dz / (T)g->dx;
        FN(pml)(D2, p[7], nz, nx, 0, m.kz, m.bz, m.az);
        for (long i = 0; i < N; ++i) ez[i] += dtdz * D2[i];
        if (src_type == 1) {
            const T lam_mu = m.L[s] + m.M[s];
            const T st = wav[it] * (T)g->dt / ((T)2 * lam_mu * (T)g->dx * (T)g->dz);
            ex[s] += st; ez[s] += st;
        }
        /* 5) exz */
        FN(deriv)(D1, U, nz, nx, 0, 1, 0, nz - 2, 1, nx - 1, c, nk);             /* Dz_U */
        FN(pml)(D1, p[8], nz, nx, 0, m.kz, m.bz, m.az);
        FN(deriv)(D2, V, nz, nx, 1, 1, 0, nx - 2, 0, nz - 2, c, nk);             /* Dx_V */
        FN(pml)(D2, p[9], nz, nx, 1, m.kx, m.bx, m.ax);
        for (int r = 0; r <= nz - 2; ++r)
            for (int j = 0; j <= nx - 2; ++j) {
                const long q = (long)r * nx + j;
                es[q] += half * (dtdz * D1[q] + dtdx * D2[q]);
            }
        /* record */
        if (rec)
            for (int i = 0; i < nrec; ++i) {
                const long q = (long)(rz_dom[i] + m.z0) * nx + rx_dom[i] + g->pad;
                rec[((long)0 * nt + it) * nrec + i] = U[q];
                rec[((long)1 * nt + it) * nrec + i] = V[q];
                rec[((long)2 * nt + it) * nrec + i] = ex[q];
                rec[((long)3 * nt + it) * nrec + i] = ez[q];
            }
        if (h_uv) { memcpy(h_uv + ((long)it * 2) * N, U, N * sizeof(T)); memcpy(h_uv + ((long)it * 2 + 1) * N, V, N * sizeof(T)); }
        if (h_e) {
            memcpy(h_e + ((long)it * 3) * N, ex, N * sizeof(T));
            memcpy(h_e + ((long)it * 3 + 1) * N, ez, N * sizeof(T));
            memcpy(h_e + ((long)it * 3 + 2) * N, es, N * sizeof(T));
        }
    }
    free(U); free(V); free(ex); free(ez); free(es);
    for (int i = 0; i < 10; ++i) free(p[i]);
    free(f1); free(f2); free(f3); free(D1); free(D2); free(D3);
    FN(model_free)(&m);
    return 0;
}

/* ------------------------------------------------------------------ */
/* adjoint (fwi/adjoint_jax_kernel.py:65-483) with optional imaging      */
/* rec_a:   4 x nt x nrec receiver adjoints (vx, vz, exx, ezz)           */
/* src_a:   nt (out) or NULL                                             */
/* imaging: 0 none; 1 Zhou (fwi/imaging.py) using h_e frame it and, when  */
/*          h_uv is given, the density term; 2 exact discrete gradient   */
/*          using h_e frame it-1, h_dxu0 and (when given) h_s            */
/* B_img:   padded buoyancy used by Zhou imaging (NULL: the model's B)   */
/* g:       4 x nz x nx padded accumulators (out, overwritten)           */
/*          Zhou:  g_lambda, g_mu, g_rho_lame                            */
/*          exact: g_(lambda+2mu), g_lambda, g_mu, g_B                   */
/* ua_hist: nt x 2 x nz x nx adjoint velocities (carry-out) or NULL      */
/* ea_hist: nt x 3 x nz x nx adjoint strains (carry-out) or NULL         */
/* ------------------------------------------------------------------ */
int FN(orc_adjoint)(const orc_geom *g, const T *vs, const T *vp, const T *rho,
                    int sx_dom, int sz_dom, int src_type, const T *wav,
                    int nrec, const int *rx_dom, const int *rz_dom,
                    const T *rec_a, T *src_a,
                    int imaging, const T *h_uv, const T *h_e, const T *h_s, const T *h_dxu0,
                    const T *B_img, T *gacc, T *ua_hist, T *ea_hist)
{
    FN(orc_model) m;
    FN(model_init)(&m, g, vs, vp, rho);
    const int nz = m.nz, nx = m.nx, nt = g->nt, fs = g->free_surface, nk = m.nk;
    const long N = (long)nz * nx;
    const double *c = m.c;
    const int sx = sx_dom + g->pad, sz = sz_dom + m.z0;
    const long s = (long)sz * nx + sx;
    T *U_a = calloc(N, sizeof(T)), *V_a = calloc(N, sizeof(T));
    T *ex_a = calloc(N, sizeof(T)), *ez_a = calloc(N, sizeof(T)), *es_a = calloc(N, sizeof(T));
    /* adjoint PML memories, same order as the forward carry */
    T *p[10];
    for (int i = 0; i < 10; ++i) p[i] = calloc(N, sizeof(T));
    T *D1 = malloc(N * sizeof(T)), *D2 = malloc(N * sizeof(T)), *D3 = malloc(N * sizeof(T));
    T *A1 = malloc(N * sizeof(T)), *A2 = malloc(N * sizeof(T)), *A3 = malloc(N * sizeof(T));
    T *A4 = malloc(N * sizeof(T)), *A5 = malloc(N * sizeof(T));
    const T dtdx = (T)(g->dt / g->dx), dtdz = (T)(g->dt / g->dz);
    const T hdtdz = (T)(0.5 * g->dt / g->dz), hdtdx = (T)(0.5 * g->dt / g->dx);
    const T *Bi = B_img ? B_img : m.B;
    T *BU = NULL, *BV = NULL, *Uap1 = NULL, *Uap2 = NULL, *Vap1 = NULL, *Vap2 = NULL;
    if (gacc) for (long i = 0; i < 4 * N; ++i) gacc[i] = 0;
    T *g0 = gacc, *g1 = gacc ? gacc + N : NULL, *g2 = gacc ? gacc + 2 * N : NULL, *g3 = gacc ? gacc + 3 * N : NULL;
    if (imaging == 1) {
        BU = malloc(N * sizeof(T)); BV = malloc(N * sizeof(T));
        Uap1 = calloc(N, sizeof(T)); Uap2 = calloc(N, sizeof(T));
        Vap1 = calloc(N, sizeof(T)); Vap2 = calloc(N, sizeof(T));
    }
    T *Dx_U_a = malloc(N * sizeof(T));
    T *dzva0 = malloc(nx * sizeof(T));

    for (int it = nt - 1; it >= 0; --it) {
        T src_t = 0;
        /* 0-T) receiver injection */
        for (int i = 0; i < nrec; ++i) {
            const long q = (long)(rz_dom[i] + m.z0) * nx + rx_dom[i] + g->pad;
            U_a[q] += rec_a[((long)0 * nt + it) * nrec + i];
            V_a[q] += rec_a[((long)1 * nt + it) * nrec + i];
            ex_a[q] += rec_a[((long)2 * nt + it) * nrec + i];
            ez_a[q] += rec_a[((long)3 * nt + it) * nrec + i];
        }
        /* 5-T) exz update */
        for (long i = 0; i < N; ++i) { D1[i] = 0; D2[i] = 0; }
        for (int r = 0; r <= nz - 2; ++r)
            for (int j = 0; j <= nx - 2; ++j) {
                const long q = (long)r * nx + j;
                D1[q] = hdtdz * es_a[q];
                D2[q] = hdtdx * es_a[q];
            }
        FN(pml_T)(D1, p[8], nz, nx, 0, m.kz, m.bz, m.az);
        FN(pml_T)(D2, p[9], nz, nx, 1, m.kx, m.bx, m.ax);
        FN(deriv_T)(U_a, D1, nz, nx, 0, 1, 0, nz - 2, 1, nx - 1, c, nk);        /* Dz_U */
        FN(deriv_T)(V_a, D2, nz, nx, 1, 1, 0, nx - 2, 0, nz - 2, c, nk);        /* Dx_V */
        /* 4.5-T) moment source */
        T exez_s = ex_a[s] + ez_a[s];
        if (src_type == 1) {
            const T lam_mu = m.L[s] + m.M[s];
            src_t = ((T)g->dt / ((T)2 * lam_mu * (T)g->dx * (T)g->dz)) * exez_s;
        }
        /* 4-T) ezz update */
        for (long i = 0; i < N; ++i) D3[i] = dtdz * ez_a[i];
        FN(pml_T)(D3, p[7], nz, nx, 0, m.kz, m.bz, m.az);
        for (int j = 0; j < nx; ++j) dzva0[j] = D3[j];   /* kept for the exact FS term */
        for (long i = 0; i < N; ++i) Dx_U_a[i] = 0;
        if (fs)
            for (int j = 0; j <= nx - 2; ++j) Dx_U_a[j] += m.Lr[j] * ((T)g->dz / (T)g->dx) * D3[j];
        FN(deriv_T)(V_a, D3, nz, nx, 0, 0, 1, nz - 1, 0, nx - 2, c, nk);        /* Dz_V */
        /* 3-T) exx update */
        for (int r = 0; r < nz; ++r)
            for (int j = 1; j < nx; ++j) Dx_U_a[(long)r * nx + j] += dtdx * ex_a[(long)r * nx + j];
        FN(pml_T)(Dx_U_a, p[6], nz, nx, 1, m.kx, m.bx, m.ax);
        FN(deriv_T)(U_a, Dx_U_a, nz, nx, 1, 0, 1, nx - 1, 0, nz - 1, c, nk);    /* Dx_U */
        /* 2-T) force source */
        if (src_type == 0) src_t = (T)(g->dt / (g->dx * g->dz)) * m.B[s] * V_a[s];
        /* 2-T) V update */
        for (long i = 0; i < N; ++i) { D1[i] = 0; D2[i] = 0; A5[i] = 0; }
        for (int r = 0; r <= nz - 2; ++r)
            for (int j = 1; j <= nx - 2; ++j) {
                const long q = (long)r * nx + j;
                D1[q] = dtdx * m.B[q] * V_a[q];
                D2[q] = dtdz * m.B[q] * V_a[q];
                A5[q] = dtdz * m.B[q] * V_a[q];
            }
        FN(pml_T)(D1, p[5], nz, nx, 1, m.kx, m.bx, m.ax);
        FN(pml_T)(D2, p[3], nz, nx, 0, m.kz, m.bz, m.az);
        FN(pml_T)(A5, p[4], nz, nx, 0, m.kz, m.bz, m.az);
        for (long i = 0; i < N; ++i) { A1[i] = 0; A2[i] = 0; A3[i] = 0; A4[i] = 0; }
        /* A1 = M_es_a, A2 = L2M_ez_a, A3 = L_ex_a */
        FN(deriv_T)(A1, D1, nz, nx, 1, 0, 1, nx - 2, 0, nz - 2, c, nk);         /* Dx_M_es */
        FN(deriv_T)(A2, D2, nz, nx, 0, 1, 0, nz - 2, 1, nx - 2, c, nk);         /* Dz_L2M_ez */
        FN(deriv_T)(A3, A5, nz, nx, 0, 1, 0, nz - 2, 1, nx - 2, c, nk);         /* Dz_L_ex */
        /* 1-T) U update */
        for (long i = 0; i < N; ++i) { D1[i] = 0; D2[i] = 0; D3[i] = 0; }
        for (int r = 0; r <= nz - 2; ++r)
            for (int j = 1; j <= nx - 2; ++j) {
                const long q = (long)r * nx + j;
                D1[q] = dtdx * m.B[q] * U_a[q];
                D2[q] = dtdx * m.B[q] * U_a[q];
                D3[q] = dtdz * m.B[q] * U_a[q];
            }
        FN(pml_T)(D1, p[0], nz, nx, 1, m.kx, m.bx, m.ax);
        FN(pml_T)(D2, p[1], nz, nx, 1, m.kx, m.bx, m.ax);
        FN(pml_T)(D3, p[2], nz, nx, 0, m.kz, m.bz, m.az);
        /* A4 = L2M_ex_a, A5 = L_ez_a */
        for (long i = 0; i < N; ++i) A5[i] = 0;
        FN(deriv_T)(A4, D1, nz, nx, 1, 1, 0, nx - 2, 0, nz - 2, c, nk);         /* Dx_L2M_ex */
        FN(deriv_T)(A5, D2, nz, nx, 1, 1, 0, nx - 2, 0, nz - 2, c, nk);         /* Dx_L_ez */
        FN(deriv_T)(A1, D3, nz, nx, 0, 0, 1, nz - 2, 1, nx - 2, c, nk);         /* Dz_M_es */
        if (fs)
            for (int j = 1; j <= nx - 2; ++j) A1[j] += (T)(2.0 * c[1]) * D3[j];

        /* exact gradient: pairs with the carry-in strains of this step */
        if (imaging == 2 && gacc) {
            const T *exm = it > 0 ? h_e + ((long)(it - 1) * 3) * N : NULL;
            const T *ezm = it > 0 ? h_e + ((long)(it - 1) * 3 + 1) * N : NULL;
            const T *esm = it > 0 ? h_e + ((long)(it - 1) * 3 + 2) * N : NULL;
            if (it > 0)
                for (long i = 0; i < N; ++i) {
                    g0[i] += exm[i] * A4[i] + ezm[i] * A2[i];
                    g1[i] += ezm[i] * A5[i] + exm[i] * A3[i];
                    g2[i] += (T)2 * esm[i] * A1[i];
                }
            if (fs && h_dxu0) {
                for (int j = 0; j <= nx - 2; ++j) {
                    const T Tt = dzva0[j] * h_dxu0[(long)it * nx + j] * (T)g->dz / (T)g->dx;
                    g1[j] += -Tt / m.L2M[j];
                    g0[j] += Tt * m.L[j] / (m.L2M[j] * m.L2M[j]);
                }
            }
            if (src_type == 1) {
                const T lam_mu = m.L[s] + m.M[s];
                const T st = wav[it] * (T)g->dt / ((T)2 * lam_mu * (T)g->dx * (T)g->dz);
                const T Q = -st / lam_mu * exez_s;
                g1[s] += Q; g2[s] += Q;
            }
            if (h_s) {
                const T *su = h_s + ((long)it * 2) * N, *sv = h_s + ((long)it * 2 + 1) * N;
                for (long i = 0; i < N; ++i) g3[i] += U_a[i] * su[i] + V_a[i] * sv[i];
            }
        }
        /* constitutive products */
        for (long i = 0; i < N; ++i) {
            ex_a[i] = ex_a[i] + m.L2M[i] * A4[i] + m.L[i] * A3[i];
            ez_a[i] = ez_a[i] + m.L[i] * A5[i] + m.L2M[i] * A2[i];
            es_a[i] = es_a[i] + (T)2 * m.M[i] * A1[i];
        }
        if (src_a) src_a[it] = src_t;
        if (ua_hist) { memcpy(ua_hist + ((long)it * 2) * N, U_a, N * sizeof(T)); memcpy(ua_hist + ((long)it * 2 + 1) * N, V_a, N * sizeof(T)); }
        if (ea_hist) {
            memcpy(ea_hist + ((long)it * 3) * N, ex_a, N * sizeof(T));
            memcpy(ea_hist + ((long)it * 3 + 1) * N, ez_a, N * sizeof(T));
            memcpy(ea_hist + ((long)it * 3 + 2) * N, es_a, N * sizeof(T));
        }

        /* Zhou imaging (fwi/imaging.py:125-152): forward frame it with the
         * adjoint carry-out of step it; centred periodic differences */
        if (imaging == 1 && gacc) {
            const T *ex = h_e + ((long)it * 3) * N, *ez = h_e + ((long)it * 3 + 1) * N, *es = h_e + ((long)it * 3 + 2) * N;
            for (long i = 0; i < N; ++i) { BU[i] = Bi[i] * U_a[i]; BV[i] = Bi[i] * V_a[i]; }
            const T i2dx = (T)(2.0 * g->dx), i2dz = (T)(2.0 * g->dz);
#pragma omp parallel for schedule(static)
            for (int r = 0; r < nz; ++r) {
                const int rp = (r + 1) % nz, rm = (r - 1 + nz) % nz;
                for (int j = 0; j < nx; ++j) {
                    const int jp = (j + 1) % nx, jm = (j - 1 + nx) % nx;
                    const long q = (long)r * nx + j;
                    const T dxBU = (BU[(long)r * nx + jp] - BU[(long)r * nx + jm]) / i2dx;
                    const T dzBV = (BV[(long)rp * nx + j] - BV[(long)rm * nx + j]) / i2dz;
                    const T dzBU = (BU[(long)rp * nx + j] - BU[(long)rm * nx + j]) / i2dz;
                    const T dxBV = (BV[(long)r * nx + jp] - BV[(long)r * nx + jm]) / i2dx;
                    g0[q] += (ex[q] + ez[q]) * (dxBU + dzBV);
                    g1[q] += (T)2 * ex[q] * dxBU + (T)2 * ez[q] * dzBV + es[q] * (dzBU + dxBV);
                }
            }
            /* density term: t = it+1 pairs v(t) with U_a(t+1) - U_a(t-1),
             * where U_a(t+1) = Uap2 and U_a(t-1) = the current carry-out */
            if (h_uv) {
                const int t = it + 1;
                if (t >= 1 && t <= nt - 2) {
                    const T *vx = h_uv + ((long)t * 2) * N, *vz = h_uv + ((long)t * 2 + 1) * N;
                    const T i2dt = (T)(2.0 * g->dt);
                    for (long i = 0; i < N; ++i)
                        g2[i] += vx[i] * ((Uap2[i] - U_a[i]) / i2dt) + vz[i] * ((Vap2[i] - V_a[i]) / i2dt);
                }
                /* shift the ring: Uap2 <- Uap1 <- U_a */
                T *tmp = Uap2; Uap2 = Uap1; Uap1 = tmp; memcpy(Uap1, U_a, N * sizeof(T));
                tmp = Vap2; Vap2 = Vap1; Vap1 = tmp; memcpy(Vap1, V_a, N * sizeof(T));
            }
        }
    }
    if (imaging == 1 && gacc) {
        const T mdt = (T)(-g->dt);
        for (long i = 0; i < 3 * N; ++i) gacc[i] *= mdt;
    }
    free(U_a); free(V_a); free(ex_a); free(ez_a); free(es_a);
    for (int i = 0; i < 10; ++i) free(p[i]);
    free(D1); free(D2); free(D3); free(A1); free(A2); free(A3); free(A4); free(A5); free(Dx_U_a); free(dzva0);
    if (BU) { free(BU); free(BV); free(Uap1); free(Uap2); free(Vap1); free(Vap2); }
    FN(model_free)(&m);
    return 0;
}

#undef CAT_
#undef CAT
#undef FN
