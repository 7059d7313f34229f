// Host side of libstfwi.so: plan lifecycle, launches and the C-ABI declared
// in include/stfwi.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdarg>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/stfwi.h"
#include "fwi_kernels.h"
#include "post_kernels.h"
#include "lbfgs_kernels.h"
#include "net_kernels.h"

using namespace stfwi;

static thread_local std::string g_err;

static int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
static int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(x)                                                                          \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess)                                                               \
            return fail(STFWI_ERR_HIP, "%s failed: %s", #x, hipGetErrorString(e_));         \
    } while (0)

#define STFWI_MAX_GROUPS 4

struct stfwi_plan {
    stfwi_config cfg;
    int esz = 4;
    int nz = 0, nx = 0, ld = 0, z0 = 0, zrows = 0, nacc = 1;
    i64 fstride = 0, origin = 0, np = 0;
    // strain history rows: physical column cp at cp + hcoff of a row of
    // hpitch elements, so that a tile's 64 columns start on a 128-byte line
    // (hcoff = pad mod 32); hnp = nz_dom * hpitch elements per field
    int hpitch = 0, hcoff = 0;
    i64 hnp = 0;
    i64 psix_len = 0, psiz_len = 0;
    // device buffers
    void *mod = nullptr;      // L2M, M, B, Bimg fields (4 x fstride) followed by Lr0 (nx)
    void *profx = nullptr, *profz = nullptr;
    double *vpmax = nullptr;
    void *state = nullptr;    // n_shots_max x 5 x fstride
    void *psix_f = nullptr, *psiz_f = nullptr;          // forward: n_shots x 5 x len
    void *psix_a[2] = {nullptr, nullptr}, *psiz_a[2] = {nullptr, nullptr};  // adjoint: n_shots x 4 x len
    void *hist = nullptr, *hist2 = nullptr, *dxu0 = nullptr;
    void *gacc = nullptr;
    int hist2_mode = 0;
    // checkpointed recomputation: every `ck` steps the forward state (5 fields
    // + C-PML memories) is saved into slot it/ck of `ckpt`; the history then
    // holds ck + 1 frames (one segment plus its carry-in) and the adjoint
    // recomputes each earlier segment from its checkpoint.  `state_a` is the
    // adjoint's own state buffer in that mode.
    int ck = 0, hist_nt = 0;
    i64 hist_cap = 0;    // history frames over all shots of one set_shots
    std::vector<int> shot_hoff;   // first history frame of each shot
    bool dry = false;    // size the allocations only (stfwi_plan_footprint)
    void *ckpt = nullptr, *state_a = nullptr;
    i64 ck_state = 0, ck_psix = 0, ck_psiz = 0;   // elements per slot region
    int tz[4] = {16, 16, 16, 16};   // tile height per time-step kernel (autotuned)
    // shots are split into `groups` contiguous groups whose time loops run on
    // their own streams, so one group's launches fill the other's tails
    int groups = 1;
    // the time-step kernels' LEAN instantiations may be used (on by default)
    int lean = 1;
    // the shot groups' time loops are enqueued by one host thread each
    int mt_launch = 1;
    hipStream_t gstream[STFWI_MAX_GROUPS] = {};
    hipEvent_t ev_fork = nullptr, ev_join[STFWI_MAX_GROUPS] = {};
    // group k + 1 starts its time loop after group k's first kernel, so
    // that concurrent groups run different half-steps (stagger_in)
    hipEvent_t ev_stag[STFWI_MAX_GROUPS] = {};
    // graph replay of the time loops: the forward and the adjoint are each
    // captured once per launch configuration (key) and replayed; their
    // inputs and outputs go through plan-owned buffers so that the captured
    // kernel arguments stay valid
    bool graphs = false;
    hipStream_t cap = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    std::vector<long long> gkey[2];
    void *rec_int = nullptr, *radj_int = nullptr, *sadj_int = nullptr, *wav_int = nullptr;
    ShotDev *shots_d = nullptr;
    i64 *rec_cell_d = nullptr, *uc_cell_d = nullptr;
    int *uc_ptr_d = nullptr, *uc_rec_d = nullptr;
    // receiver adjoints folded into the adjoint time-step kernels: a map of
    // every cell to its merged receiver index (-1 if none), per-row prefix
    // counts of merged receiver cells, and the per-step merged sums
    int *cellmap_d = nullptr, *rowcum_d = nullptr;
    double *lpart = nullptr;   // [shot][time blocks][nrec_max] sums of squares of the residual
    // adjoint energy accumulator of the next stfwi_adjoint (caller-owned)
    void *aen = nullptr;
    int aen_comp = 0;
    void *injm = nullptr;
    std::vector<int> shot_nt;
    i64 bytes = 0;
    // shots
    int n_shots = 0, nt_run = 0, max_nuc = 0, max_nrec = 0;
    const void *wav = nullptr;
    bool have_model = false, fwd_done = false, adj_done = false;
    std::vector<void *> allocs;
};

static int dalloc(stfwi_plan *p, void **ptr, size_t bytes)
{
    if (bytes == 0) bytes = 256;
    if (p->dry) {
        *ptr = nullptr;
        p->bytes += (i64)bytes;
        return 0;
    }
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
        *ptr = nullptr;
        return fail(STFWI_ERR_OOM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    p->allocs.push_back(*ptr);
    p->bytes += (i64)bytes;
    return 0;
}

template <typename T>
static void fill(void *ptr, i64 n, hipStream_t st)
{
    if (n <= 0) return;
    int blocks = (int)std::min<i64>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_fill<T>, dim3(blocks), dim3(256), 0, st, (T *)ptr, n, T(0));
}

static int popcount5(int m)
{
    int n = 0;
    for (int f = 0; f < 5; ++f) n += (m >> f) & 1;
    return n;
}

template <typename T>
static Geo<T> make_geo(const stfwi_plan *p)
{
    const stfwi_config &c = p->cfg;
    Geo<T> g;
    g.nz = p->nz; g.nx = p->nx; g.ld = p->ld;
    g.z0 = p->z0; g.pad = c.pad; g.nzd = c.nz_dom; g.nxd = c.nx_dom;
    g.fs = c.free_surface; g.zrows = p->zrows;
    g.fstride = p->fstride; g.origin = p->origin; g.np = p->np;
    g.hpitch = p->hpitch; g.hcoff = p->hcoff; g.hnp = p->hnp;
    g.dt = (T)c.dt; g.dx = (T)c.dx; g.dz = (T)c.dz;
    g.dtdx = (T)(c.dt / c.dx); g.dtdz = (T)(c.dt / c.dz);
    g.hdtdx = (T)(0.5 * c.dt / c.dx); g.hdtdz = (T)(0.5 * c.dt / c.dz);
    g.dzdx = (T)(c.dz / c.dx);
    g.fsrc = (T)(c.dt / (c.dx * c.dz));
    return g;
}

extern "C" int stfwi_abi_version(void) { return STFWI_ABI_VERSION; }
extern "C" const char *stfwi_last_error(void) { return g_err.c_str(); }

extern "C" long long stfwi_plan_bytes(const stfwi_plan *plan) { return plan ? plan->bytes : 0; }

extern "C" int stfwi_plan_destroy(stfwi_plan *plan)
{
    if (!plan) return 0;
    // work of the plan may still be queued: finish it before freeing
    if (plan->cap) (void)hipStreamSynchronize(plan->cap);
    for (int k = 0; k < STFWI_MAX_GROUPS; ++k)
        if (plan->gstream[k]) (void)hipStreamSynchronize(plan->gstream[k]);
    for (void *a : plan->allocs) (void)hipFree(a);
    for (int k = 0; k < STFWI_MAX_GROUPS; ++k) {
        if (plan->gstream[k]) (void)hipStreamDestroy(plan->gstream[k]);
        if (plan->ev_join[k]) (void)hipEventDestroy(plan->ev_join[k]);
        if (plan->ev_stag[k]) (void)hipEventDestroy(plan->ev_stag[k]);
    }
    if (plan->ev_fork) (void)hipEventDestroy(plan->ev_fork);
    for (int k = 0; k < 2; ++k)
        if (plan->gexec[k]) (void)hipGraphExecDestroy(plan->gexec[k]);
    if (plan->cap) (void)hipStreamDestroy(plan->cap);
    if (plan->ev_in) (void)hipEventDestroy(plan->ev_in);
    if (plan->ev_out) (void)hipEventDestroy(plan->ev_out);
    delete plan;
    return 0;
}

static int validate_config(const stfwi_config &c)
{
    if (c.fd_order != 2 && c.fd_order != 4 && c.fd_order != 8)
        return fail(STFWI_ERR_UNSUPPORTED, "fd_order=%d not supported (use 2, 4, or 8)", c.fd_order);
    if (c.dtype != STFWI_F32 && c.dtype != STFWI_F64) return fail(STFWI_ERR_INVALID, "dtype=%d", c.dtype);
    if (c.nz_dom < 1 || c.nx_dom < 1 || c.pad < 1 || c.n_shots_max < 1 || c.nt_max < 1 || c.nrec_max < 0)
        return fail(STFWI_ERR_INVALID, "bad sizes");
    if (c.imaging < 0 || c.imaging > 2 || c.grad_params < 1 || c.grad_params > 3)
        return fail(STFWI_ERR_INVALID, "imaging=%d grad_params=%d", c.imaging, c.grad_params);
    if (c.dx <= 0 || c.dz <= 0 || c.dt <= 0) return fail(STFWI_ERR_INVALID, "dx, dz, dt must be positive");
    if (c.ckpt_every < 0) return fail(STFWI_ERR_INVALID, "ckpt_every=%d", c.ckpt_every);
    if (c.hist_frames < 0) return fail(STFWI_ERR_INVALID, "hist_frames=%lld", c.hist_frames);
    if (c.ckpt_every > 0 && c.imaging == STFWI_IMAGING_ZHOU && c.grad_params == STFWI_GRAD_ALL)
        return fail(STFWI_ERR_UNSUPPORTED, "checkpointing does not support Zhou imaging of rho (it pairs steps "
                                           "across segment ends)");
    const int nz = c.nz_dom + (c.free_surface ? c.pad : 2 * c.pad), nx = c.nx_dom + 2 * c.pad;
    if (nz < 10 || nx < 10) return fail(STFWI_ERR_INVALID, "padded grid %dx%d too small", nz, nx);
    return 0;
}

// geometry and every device allocation of a plan (sizes only when p->dry)
static int plan_alloc(stfwi_plan *p)
{
    const stfwi_config &c = p->cfg;
    p->esz = c.dtype == STFWI_F64 ? 8 : 4;
    for (int k = 0; k < 4; ++k) p->tz[k] = c.dtype == STFWI_F64 ? 8 : 16;
    p->nz = c.nz_dom + (c.free_surface ? c.pad : 2 * c.pad);
    p->nx = c.nx_dom + 2 * c.pad;
    p->z0 = c.free_surface ? 0 : c.pad;
    p->zrows = c.free_surface ? c.pad : 2 * c.pad;
    p->ld = ((p->nx + 8 + 63) / 64) * 64;
    p->fstride = (i64)(p->nz + 8) * p->ld;
    p->origin = (i64)4 * p->ld;
    p->np = (i64)c.nz_dom * c.nx_dom;
    p->hcoff = c.pad % 32;
    p->hpitch = (c.nx_dom + p->hcoff + 31) / 32 * 32;
    p->hnp = (i64)c.nz_dom * p->hpitch;
    p->psix_len = (i64)p->nz * 2 * c.pad;
    p->psiz_len = (i64)p->zrows * p->ld;
    p->nacc = c.grad_params == 3 ? (c.imaging == 1 ? 3 : 4) : c.grad_params;
    const size_t es = p->esz;
    const i64 S = c.n_shots_max;
    int rc = 0;
    rc |= dalloc(p, &p->mod, (4 * p->fstride + p->nx) * es);
    rc |= dalloc(p, &p->profx, 5 * p->nx * es);
    rc |= dalloc(p, &p->profz, 5 * p->nz * es);
    rc |= dalloc(p, (void **)&p->vpmax, sizeof(double));
    rc |= dalloc(p, &p->state, S * 5 * p->fstride * es);
    rc |= dalloc(p, &p->psix_f, S * 5 * p->psix_len * es);
    rc |= dalloc(p, &p->psiz_f, S * 5 * p->psiz_len * es);
    for (int k = 0; k < 2; ++k) {
        rc |= dalloc(p, &p->psix_a[k], S * 4 * p->psix_len * es);
        rc |= dalloc(p, &p->psiz_a[k], S * 4 * p->psiz_len * es);
    }
    if (c.store_history && c.imaging) {
        p->ck = c.ckpt_every > 0 && c.ckpt_every < c.nt_max ? c.ckpt_every : 0;
        p->hist_nt = p->ck ? p->ck + 1 : c.nt_max;
        p->hist_cap = p->ck || c.hist_frames == 0 ? S * p->hist_nt : c.hist_frames;
        rc |= dalloc(p, &p->hist, (size_t)p->hist_cap * 3 * p->hnp * es);
        if (c.grad_params == 3) {
            p->hist2_mode = c.imaging == 2 ? 1 : 2;
            rc |= dalloc(p, &p->hist2, (size_t)p->hist_cap * 2 * p->np * es);
        }
        if (p->ck) {
            const i64 nslot = (c.nt_max + p->ck - 1) / p->ck;
            p->ck_state = S * 5 * p->fstride;
            p->ck_psix = S * 5 * p->psix_len;
            p->ck_psiz = S * 5 * p->psiz_len;
            rc |= dalloc(p, &p->ckpt, (size_t)nslot * (p->ck_state + p->ck_psix + p->ck_psiz) * es);
            rc |= dalloc(p, &p->state_a, S * 5 * p->fstride * es);
        }
        if (c.imaging == 2 && c.free_surface) rc |= dalloc(p, &p->dxu0, (size_t)S * c.nt_max * p->nx * es);
        rc |= dalloc(p, &p->gacc, (size_t)S * p->nacc * p->hnp * es);
    }
    rc |= dalloc(p, (void **)&p->shots_d, S * sizeof(ShotDev));
    rc |= dalloc(p, &p->rec_int, (size_t)S * 4 * c.nt_max * std::max(c.nrec_max, 1) * es);
    rc |= dalloc(p, &p->radj_int, (size_t)S * 4 * c.nt_max * std::max(c.nrec_max, 1) * es);
    rc |= dalloc(p, &p->sadj_int, (size_t)S * c.nt_max * es);
    rc |= dalloc(p, &p->wav_int, (size_t)S * c.nt_max * es);
    rc |= dalloc(p, (void **)&p->rec_cell_d, (S * std::max(c.nrec_max, 1)) * sizeof(i64));
    rc |= dalloc(p, (void **)&p->uc_cell_d, (S * std::max(c.nrec_max, 1)) * sizeof(i64));
    rc |= dalloc(p, (void **)&p->uc_ptr_d, (S * std::max(c.nrec_max, 1) + 1) * sizeof(int));
    rc |= dalloc(p, (void **)&p->uc_rec_d, (S * std::max(c.nrec_max, 1)) * sizeof(int));
    rc |= dalloc(p, (void **)&p->cellmap_d, S * p->fstride * sizeof(int));
    rc |= dalloc(p, (void **)&p->rowcum_d, S * (p->nz + 1) * sizeof(int));
    rc |= dalloc(p, (void **)&p->lpart,
                 S * std::max(c.nrec_max, 1) * ((std::max(c.nt_max, 1) + RES_TCH - 1) / RES_TCH) * sizeof(double));
    rc |= dalloc(p, &p->injm, (size_t)S * c.nt_max * 4 * std::max(c.nrec_max, 1) * es);
    return rc;
}

extern "C" int stfwi_plan_footprint(const stfwi_config *cfg, long long *bytes)
{
    if (!cfg || !bytes) return fail(STFWI_ERR_INVALID, "null argument");
    const int v = validate_config(*cfg);
    if (v) return v;
    stfwi_plan p;
    p.cfg = *cfg;
    p.dry = true;
    plan_alloc(&p);
    *bytes = p.bytes;
    return 0;
}

extern "C" int stfwi_plan_choose_ckpt(const stfwi_config *cfg, long long budget_bytes, int *ckpt_every)
{
    if (!cfg || !ckpt_every) return fail(STFWI_ERR_INVALID, "null argument");
    stfwi_config c = *cfg;
    c.ckpt_every = 0;
    int v = validate_config(c);
    if (v) return v;
    long long full = 0;
    stfwi_plan_footprint(&c, &full);
    if (full <= budget_bytes || !c.store_history || !c.imaging) {
        *ckpt_every = 0;
        return 0;
    }
    if (c.imaging == STFWI_IMAGING_ZHOU && c.grad_params == STFWI_GRAD_ALL)
        return fail(STFWI_ERR_UNSUPPORTED, "the history needs %lld bytes of %lld and Zhou imaging of rho cannot "
                                           "be checkpointed", full, budget_bytes);
    // footprint(K) = nslot(K) states + (K + 1) history frames, least near
    // K = sqrt(nt * state bytes / frame bytes): scan every K, the set-up is
    // host arithmetic only
    long long best = -1;
    int best_k = 0;
    for (int k = 1; k < c.nt_max; ++k) {
        c.ckpt_every = k;
        long long b = 0;
        stfwi_plan_footprint(&c, &b);
        if (best < 0 || b < best) {
            best = b;
            best_k = k;
        }
    }
    if (best < 0 || best > budget_bytes)
        return fail(STFWI_ERR_OOM, "no checkpoint spacing fits: the smallest plan needs %lld bytes of %lld", best,
                    budget_bytes);
    *ckpt_every = best_k;
    return 0;
}

extern "C" int stfwi_plan_create(const stfwi_config *cfg, stfwi_plan **out)
{
    if (!cfg || !out) return fail(STFWI_ERR_INVALID, "null argument");
    *out = nullptr;
    const int v = validate_config(*cfg);
    if (v) return v;
    const stfwi_config &c = *cfg;
    stfwi_plan *p = new stfwi_plan();
    p->cfg = c;
    const int rc = plan_alloc(p);
    const i64 S = c.n_shots_max;
    if (rc) {
        std::string e = g_err;
        stfwi_plan_destroy(p);
        g_err = e;
        return STFWI_ERR_OOM;
    }
    // streams of the shot groups (default priority) and their fork/join events
    for (int k = 0; k < STFWI_MAX_GROUPS; ++k) {
        HIP_TRY(hipStreamCreateWithFlags(&p->gstream[k], hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&p->ev_join[k], hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&p->ev_stag[k], hipEventDisableTiming));
    }
    HIP_TRY(hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming));
    HIP_TRY(hipStreamCreateWithFlags(&p->cap, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&p->ev_in, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&p->ev_out, hipEventDisableTiming));
    // margins of the moduli buffers must be zero
    hipStream_t st = nullptr;
    if (c.dtype == STFWI_F64) fill<double>(p->mod, 4 * p->fstride + p->nx, st);
    else fill<float>(p->mod, 4 * p->fstride + p->nx, st);
    // no receivers until stfwi_set_shots
    fill<int>(p->rowcum_d, S * (p->nz + 1), st);
    HIP_TRY(hipStreamSynchronize(st));
    *out = p;
    return 0;
}

template <typename T>
static int set_model_t(stfwi_plan *p, const void *vs, const void *vp, const void *rho, const void *b_img,
                       hipStream_t st)
{
    Geo<T> g = make_geo<T>(p);
    T *mod = (T *)p->mod;
    T *L2M = mod, *M = mod + p->fstride, *B = mod + 2 * p->fstride, *Bimg = mod + 3 * p->fstride;
    T *Lr0 = mod + 4 * p->fstride;
    dim3 grid((p->nx + 127) / 128, p->nz);
    hipLaunchKernelGGL(k_model<T>, grid, dim3(128), 0, st, g, (const T *)vs, (const T *)vp, (const T *)rho,
                       (const T *)b_img, L2M, M, B, Bimg, Lr0);
    hipLaunchKernelGGL(k_vpmax<T>, dim3(1), dim3(256), 0, st, (const T *)vp, p->np, p->vpmax);
    const int n = std::max(p->nx, p->nz);
    hipLaunchKernelGGL(k_profiles<T>, dim3((n + 127) / 128), dim3(128), 0, st, p->nz, p->nx, p->cfg.pad,
                       p->cfg.free_surface, p->cfg.nz_dom, p->cfg.dx, p->cfg.dt, p->cfg.fc, p->vpmax,
                       (T *)p->profx, (T *)p->profz);
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int stfwi_set_model(stfwi_plan *p, const void *vs, const void *vp, const void *rho, const void *b_img,
                               void *stream)
{
    if (!p || !vs || !vp || !rho) return fail(STFWI_ERR_INVALID, "null argument");
    int rc = p->cfg.dtype == STFWI_F64 ? set_model_t<double>(p, vs, vp, rho, b_img, (hipStream_t)stream)
                                       : set_model_t<float>(p, vs, vp, rho, b_img, (hipStream_t)stream);
    if (rc == 0) p->have_model = true;
    p->fwd_done = false;
    return rc;
}

extern "C" int stfwi_set_shots(stfwi_plan *p, const stfwi_shot *shots, int n_shots, const int *rec_x,
                               const int *rec_z, const void *wavelets, void *stream)
{
    if (!p || !shots || !wavelets) return fail(STFWI_ERR_INVALID, "null argument");
    const stfwi_config &c = p->cfg;
    if (n_shots < 1 || n_shots > c.n_shots_max)
        return fail(STFWI_ERR_INVALID, "n_shots=%d outside [1, %d]", n_shots, c.n_shots_max);
    std::vector<ShotDev> sd(n_shots);
    std::vector<i64> cells(n_shots * (size_t)std::max(c.nrec_max, 1));
    std::vector<i64> ucell;
    std::vector<int> uptr(1, 0), urec;
    std::vector<int> rowcum((size_t)n_shots * (p->nz + 1), 0), shot_nt(n_shots);
    int nt_run = 0, max_nuc = 0, max_nrec = 0;
    std::vector<int> hoff(n_shots, 0);
    i64 frames = 0;
    for (int s = 0; s < n_shots; ++s) {
        // each shot's history holds its own record (or one checkpoint
        // segment), packed one shot after another
        const int nf = p->ck ? p->hist_nt : shots[s].nt;
        if (p->hist && frames + std::max(nf, 0) > p->hist_cap)
            return fail(STFWI_ERR_INVALID, "the shots' history needs more than the plan's %lld frames (shot %d)",
                        (long long)p->hist_cap, s);
        hoff[s] = (int)frames;
        frames += std::max(nf, 0);
    }
    for (int s = 0; s < n_shots; ++s) {
        const stfwi_shot &h = shots[s];
        if (h.src_type != STFWI_SRC_FORCE && h.src_type != STFWI_SRC_MOMENT)
            return fail(STFWI_ERR_UNSUPPORTED, "src_type=%d not supported (use force or moment)", h.src_type);
        if (h.nt < 1 || h.nt > c.nt_max) return fail(STFWI_ERR_INVALID, "shot %d: nt=%d outside [1, %d]", s, h.nt, c.nt_max);
        if (h.nrec < 1 || h.nrec > c.nrec_max)
            return fail(STFWI_ERR_INVALID, "shot %d: nrec=%d outside [1, %d]", s, h.nrec, c.nrec_max);
        if (h.src_x < 0 || h.src_x >= c.nx_dom || h.src_z < 0 || h.src_z >= c.nz_dom)
            return fail(STFWI_ERR_INVALID, "shot %d: source (%d,%d) outside the model", s, h.src_z, h.src_x);
        ShotDev &d = sd[s];
        d.sx = h.src_x + c.pad;
        d.sz = h.src_z + p->z0;
        d.type = h.src_type;
        d.nt = h.nt;
        d.hoff = hoff[s];
        d.rec_off = s * c.nrec_max;
        d.nrec = h.nrec;
        d.uc_off = (int)ucell.size();
        std::map<i64, std::vector<int>> byc;
        for (int j = 0; j < h.nrec; ++j) {
            const int rx = rec_x[h.rec_offset + j], rz = rec_z[h.rec_offset + j];
            if (rx < 0 || rx >= c.nx_dom || rz < 0 || rz >= c.nz_dom)
                return fail(STFWI_ERR_INVALID, "shot %d receiver %d: (%d,%d) outside the model", s, j, rz, rx);
            const i64 cell = (i64)(rz + p->z0) * p->ld + rx + c.pad;
            cells[(size_t)s * c.nrec_max + j] = cell;
            byc[cell].push_back(j);
        }
        int *rcum = rowcum.data() + (size_t)s * (p->nz + 1);
        for (auto &kv : byc) {
            ucell.push_back(kv.first);
            for (int j : kv.second) urec.push_back(j);
            uptr.push_back((int)urec.size());
            ++rcum[kv.first / p->ld + 1];
        }
        for (int r = 0; r < p->nz; ++r) rcum[r + 1] += rcum[r];
        d.nuc = (int)ucell.size() - d.uc_off;
        shot_nt[s] = h.nt;
        nt_run = std::max(nt_run, h.nt);
        max_nuc = std::max(max_nuc, d.nuc);
        max_nrec = std::max(max_nrec, h.nrec);
    }
    hipStream_t st = (hipStream_t)stream;
    HIP_TRY(hipMemcpyAsync(p->shots_d, sd.data(), sd.size() * sizeof(ShotDev), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(p->rec_cell_d, cells.data(), cells.size() * sizeof(i64), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(p->uc_cell_d, ucell.data(), ucell.size() * sizeof(i64), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(p->uc_ptr_d, uptr.data(), uptr.size() * sizeof(int), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(p->uc_rec_d, urec.data(), urec.size() * sizeof(int), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(p->rowcum_d, rowcum.data(), rowcum.size() * sizeof(int), hipMemcpyHostToDevice, st));
    {
        const i64 n = (i64)n_shots * p->fstride;
        hipLaunchKernelGGL(k_fill<int>, dim3((int)std::min<i64>((n + 255) / 256, 65536)), dim3(256), 0, st,
                           p->cellmap_d, n, -1);
        hipLaunchKernelGGL(k_cellmap, dim3((max_nuc + 127) / 128, n_shots), dim3(128), 0, st,
                           (const ShotDev *)p->shots_d, p->uc_cell_d, p->fstride, p->origin, p->cellmap_d);
        HIP_TRY(hipGetLastError());
    }
    // the wavelets are read by the kernels of every later time loop: keep a copy
    HIP_TRY(hipMemcpyAsync(p->wav_int, wavelets, (size_t)n_shots * c.nt_max * p->esz, hipMemcpyDeviceToDevice, st));
    // the host vectors die at return: wait for the copies
    HIP_TRY(hipStreamSynchronize(st));
    p->n_shots = n_shots;
    p->nt_run = nt_run;
    p->max_nuc = max_nuc;
    p->max_nrec = max_nrec;
    p->shot_nt = shot_nt;
    p->shot_hoff = hoff;
    p->wav = p->wav_int;
    p->fwd_done = false;
    return 0;
}


// tile heights compiled per precision: 4/8/16/32 rows in FP32, 4/8 in FP64
// (larger FP64 tiles would exceed 64 KiB of LDS in the forward kernel)
template <typename T>
static bool tz_supported(int tz)
{
    return std::is_same<T, float>::value ? (tz == 4 || tz == 8 || tz == 16 || tz == 32) : (tz == 4 || tz == 8);
}

// kernel ids: 0 forward velocity, 1 forward strain, 2 adjoint strain,
// 3 adjoint velocity
template <typename T, int FD>
static void launch_fwd(int which, int tz, FwdArgs<T> a, int rblk, int S, hipStream_t st, int it)
{
    a.tiles_z = (a.g.nz + tz - 1) / tz;
    a.ntiles = a.tiles_x * a.tiles_z;
    a.dgx = make_udiv((unsigned)(a.ntiles + rblk));
    a.dtx = make_udiv((unsigned)a.tiles_x);
    dim3 grid(a.ntiles + rblk, 1, S), blk(TILE_X, BLOCK_Y);
    // the LEAN instantiations fix what these arguments already hold
    const bool lean = a.lean && a.hist2_mode == 0 && (which == 0 || (!a.illum && !a.snap));
#define STFWI_LF(TZV)                                                                              \
    do {                                                                                           \
        if (which == 0) {                                                                          \
            if (lean) hipLaunchKernelGGL((k_fwd_vel<T, FD, TZV, true>), grid, blk, 0, st, a, it);     \
            else hipLaunchKernelGGL((k_fwd_vel<T, FD, TZV, false>), grid, blk, 0, st, a, it);         \
        } else {                                                                                   \
            if (lean) hipLaunchKernelGGL((k_fwd_strain<T, FD, TZV, true>), grid, blk, 0, st, a, it);  \
            else hipLaunchKernelGGL((k_fwd_strain<T, FD, TZV, false>), grid, blk, 0, st, a, it);      \
        }                                                                                          \
    } while (0)
    if constexpr (std::is_same<T, float>::value) {
        if (tz == 32) STFWI_LF(32);
        else if (tz == 16) STFWI_LF(16);
        else if (tz == 4) STFWI_LF(4);
        else STFWI_LF(8);
    } else {
        if (tz == 4) STFWI_LF(4);
        else STFWI_LF(8);
    }
#undef STFWI_LF
}

template <typename T, int FD>
static void launch_adj(int which, int tz, AdjArgs<T> a, int S, hipStream_t st, int it)
{
    a.tiles_z = (a.g.nz + tz - 1) / tz;
    a.ntiles = a.tiles_x * a.tiles_z;
    a.dgx = make_udiv((unsigned)a.ntiles);
    a.dtx = make_udiv((unsigned)a.tiles_x);
    dim3 grid(a.ntiles, 1, S), blk(TILE_X, BLOCK_Y);
    // the LEAN instantiations fix what these arguments already hold
    const bool lean = a.lean && (which == 2 ? !a.src_adj
                                            : (!a.snap && !a.aen && a.imaging == 2 && a.gparams == 1));
#define STFWI_LA(TZV)                                                                              \
    do {                                                                                           \
        if (which == 2) {                                                                          \
            if (lean) hipLaunchKernelGGL((k_adj_strain<T, FD, TZV, true>), grid, blk, 0, st, a, it);  \
            else hipLaunchKernelGGL((k_adj_strain<T, FD, TZV, false>), grid, blk, 0, st, a, it);      \
        } else {                                                                                   \
            if (lean) hipLaunchKernelGGL((k_adj_vel<T, FD, TZV, true>), grid, blk, 0, st, a, it);     \
            else hipLaunchKernelGGL((k_adj_vel<T, FD, TZV, false>), grid, blk, 0, st, a, it);         \
        }                                                                                          \
    } while (0)
    if constexpr (std::is_same<T, float>::value) {
        if (tz == 32) STFWI_LA(32);
        else if (tz == 16) STFWI_LA(16);
        else if (tz == 4) STFWI_LA(4);
        else STFWI_LA(8);
    } else {
        if (tz == 4) STFWI_LA(4);
        else STFWI_LA(8);
    }
#undef STFWI_LA
}

// contiguous shot groups; returns the group count, g0[k]..g0[k+1] bound group k
static int group_bounds(const stfwi_plan *p, int *g0)
{
    const int G = std::max(1, std::min(p->groups, p->n_shots));
    for (int k = 0; k <= G; ++k) g0[k] = (int)((long long)p->n_shots * k / G);
    return G;
}

// shortest record among shots [s0, s1): the time-step kernels of a launch
// test each block's shot for a finished record only from that step on
static int nt_min_of(const stfwi_plan *p, int s0, int s1)
{
    int m = INT_MAX;
    for (int s = s0; s < s1; ++s) m = std::min(m, p->shot_nt[s]);
    return m;
}

static hipError_t fork_groups(stfwi_plan *p, int G, hipStream_t st, hipStream_t *gs)
{
    if (G == 1) {
        gs[0] = st;
        return hipSuccess;
    }
    hipError_t e = hipEventRecord(p->ev_fork, st);
    for (int k = 0; k < G && e == hipSuccess; ++k) {
        gs[k] = p->gstream[k];
        e = hipStreamWaitEvent(gs[k], p->ev_fork, 0);
    }
    return e;
}

// Staggered start of the shot groups' time loops: group k waits for the
// first kernel of group k - 1 (stagger_in before group k's first launch,
// stagger_out after group k's first kernel).  The groups then run offset by
// a half-step each, so that concurrent launches are of different kernels
// and fill each other's ramps and tails; the order of every group's own
// launches is unchanged, so results are the same bit for bit.
static hipError_t stagger_in(stfwi_plan *p, int k, hipStream_t *gs)
{
    return k > 0 ? hipStreamWaitEvent(gs[k], p->ev_stag[k - 1], 0) : hipSuccess;
}

static hipError_t stagger_out(stfwi_plan *p, int k, int G, hipStream_t *gs)
{
    return k + 1 < G ? hipEventRecord(p->ev_stag[k], gs[k]) : hipSuccess;
}

static hipError_t join_groups(stfwi_plan *p, int G, hipStream_t st)
{
    if (G == 1) return hipSuccess;
    hipError_t e = hipSuccess;
    for (int k = 0; k < G && e == hipSuccess; ++k) {
        e = hipEventRecord(p->ev_join[k], p->gstream[k]);
        if (e == hipSuccess) e = hipStreamWaitEvent(st, p->ev_join[k], 0);
    }
    return e;
}

// The time loops of the shot groups are enqueued by one host thread per
// group (group 0 on the calling thread), so that the host's launch rate
// does not bound the loops: every stream gets the same launches in the same
// order as from one thread.  body(k) enqueues group k's work; the calling
// thread's device is made current in each worker; the first failing
// group's status and message are returned.  Threads are used only when the
// plan allows it and the caller's stream is not being captured.
struct GroupHandoff {
    // group k has enqueued its first kernel (and the event that group k + 1
    // waits on), or has stopped
    std::atomic<int> ready[STFWI_MAX_GROUPS];
    GroupHandoff()
    {
        for (auto &r : ready) r.store(0, std::memory_order_relaxed);
    }
    void wait_prev(int k)
    {
        if (k > 0)
            while (!ready[k - 1].load(std::memory_order_acquire)) std::this_thread::yield();
    }
    void mark(int k) { ready[k].store(1, std::memory_order_release); }
};

template <typename F>
static int for_groups(stfwi_plan *p, int G, hipStream_t st, F body)
{
    GroupHandoff h;
    auto run = [&](int k) {
        const int rc = body(k, h);
        h.mark(k);          // also when body stopped early: no group waits forever
        return rc;
    };
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    bool capturing = false;
    if (G > 1) {
        if (hipStreamIsCapturing(st, &cs) == hipSuccess) capturing = cs != hipStreamCaptureStatusNone;
        else (void)hipGetLastError();   // not a capture: leave no error behind
    }
    const bool threads = G > 1 && p->mt_launch && !capturing;
    if (!threads) {
        for (int k = 0; k < G; ++k) {
            const int rc = run(k);
            if (rc) return rc;
        }
        return 0;
    }
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    int rcs[STFWI_MAX_GROUPS] = {};
    std::string msg[STFWI_MAX_GROUPS];
    std::thread th[STFWI_MAX_GROUPS];
    int spawned = 1;
    for (int k = 1; k < G; ++k) {
        try {
            th[k] = std::thread([&, k] {
                const hipError_t e = hipSetDevice(dev);
                rcs[k] = e == hipSuccess ? run(k) : fail(STFWI_ERR_HIP, "hipSetDevice failed: %s", hipGetErrorString(e));
                if (e != hipSuccess) h.mark(k);
                if (rcs[k]) msg[k] = g_err;
            });
        } catch (...) {
            break;          // no thread: the remaining groups run here below
        }
        spawned = k + 1;
    }
    rcs[0] = run(0);
    for (int k = spawned; k < G; ++k) rcs[k] = run(k);
    for (int k = 1; k < spawned; ++k) th[k].join();
    for (int k = 0; k < G; ++k)
        if (rcs[k]) {
            if (k > 0 && k < spawned) g_err = msg[k];
            return rcs[k];
        }
    return 0;
}

template <typename T>
static FwdArgs<T> fwd_args(stfwi_plan *p, void *rec, int rec_mask, void *illum, int illum_comp, void *snap,
                           int snap_mask)
{
    const stfwi_config &c = p->cfg;
    FwdArgs<T> a;
    memset(&a, 0, sizeof a);
    a.g = make_geo<T>(p);
    a.shots = p->shots_d;
    const T *mod = (const T *)p->mod;
    a.L2M = mod; a.M = mod + p->fstride; a.B = mod + 2 * p->fstride; a.Lr0 = mod + 4 * p->fstride;
    a.state = (T *)p->state; a.sstride = 5 * p->fstride;
    a.psix = (T *)p->psix_f; a.psiz = (T *)p->psiz_f;
    a.psix_n = p->psix_len; a.psiz_n = p->psiz_len;
    a.psix_stride = 5 * p->psix_len; a.psiz_stride = 5 * p->psiz_len;
    a.profx = (const T *)p->profx; a.profz = (const T *)p->profz;
    a.wav = (const T *)p->wav; a.nt_max = c.nt_max;
    a.rec = (T *)rec; a.nrec_max = c.nrec_max; a.rec_mask = rec_mask;
    a.rec_cell = p->rec_cell_d;
    a.hist = (T *)p->hist;
    a.hist2 = (T *)p->hist2; a.hist2_mode = p->hist ? p->hist2_mode : 0;
    a.hist_nt = p->hist_nt; a.hist_t0 = 0;
    a.dxu0 = (T *)p->dxu0;
    a.illum = (T *)illum; a.illum_comp = illum_comp;
    a.snap = (T *)snap; a.snap_mask = snap_mask; a.nsnap = popcount5(snap_mask);
    a.tiles_x = (p->nx + TILE_X - 1) / TILE_X;
    a.lean = p->lean;
    return a;
}

template <typename T>
static AdjArgs<T> adj_args(stfwi_plan *p, const void *rec_adj, void *src_adj, void *snap, int snap_mask)
{
    const stfwi_config &c = p->cfg;
    AdjArgs<T> a;
    memset(&a, 0, sizeof a);
    a.g = make_geo<T>(p);
    a.shots = p->shots_d;
    const T *mod = (const T *)p->mod;
    a.L2M = mod; a.M = mod + p->fstride; a.B = mod + 2 * p->fstride; a.Bimg = mod + 3 * p->fstride;
    a.Lr0 = mod + 4 * p->fstride;
    a.state = (T *)p->state; a.sstride = 5 * p->fstride;
    a.psix_len = p->psix_len; a.psiz_len = p->psiz_len;
    a.psix_stride = 4 * p->psix_len; a.psiz_stride = 4 * p->psiz_len;
    a.profx = (const T *)p->profx; a.profz = (const T *)p->profz;
    a.wav = (const T *)p->wav; a.nt_max = c.nt_max;
    a.rec_adj = (const T *)rec_adj; a.nrec_max = c.nrec_max;
    a.uc_cell = p->uc_cell_d; a.uc_ptr = p->uc_ptr_d; a.uc_rec = p->uc_rec_d;
    a.cellmap = p->cellmap_d; a.rowcum = p->rowcum_d; a.injm = (const T *)p->injm;
    a.nuc_max = std::max(c.nrec_max, 1);
    a.src_adj = (T *)src_adj;
    a.imaging = p->hist ? c.imaging : 0;
    a.gparams = c.grad_params; a.nacc = p->nacc;
    a.hist = (const T *)p->hist; a.hist2 = (const T *)p->hist2; a.dxu0 = (const T *)p->dxu0;
    a.hist_nt = p->hist_nt; a.hist_t0 = 0;
    if (p->ck) a.state = (T *)p->state_a;
    a.gacc = (T *)p->gacc;
    a.snap = (T *)snap; a.snap_mask = snap_mask; a.nsnap = popcount5(snap_mask);
    a.aen = (T *)p->aen; a.aen_comp = p->aen_comp;
    a.tiles_x = (p->nx + TILE_X - 1) / TILE_X;
    a.lean = p->lean;
    return a;
}

// checkpoint slot k <- (restore = false) or -> (restore = true) the forward
// state and C-PML memories of shots [s0, s1), on stream st
template <typename T>
static int ckpt_copy(stfwi_plan *p, int k, int s0, int s1, bool restore, hipStream_t st)
{
    const size_t es = sizeof(T);
    char *slot = (char *)p->ckpt + (size_t)k * (p->ck_state + p->ck_psix + p->ck_psiz) * es;
    struct R { char *buf; char *ck; i64 per_shot; } r[3] = {
        {(char *)p->state, slot, 5 * p->fstride},
        {(char *)p->psix_f, slot + p->ck_state * es, 5 * p->psix_len},
        {(char *)p->psiz_f, slot + (p->ck_state + p->ck_psix) * es, 5 * p->psiz_len}};
    for (auto &x : r) {
        const size_t off = (size_t)s0 * x.per_shot * es, n = (size_t)(s1 - s0) * x.per_shot * es;
        if (restore) HIP_TRY(hipMemcpyAsync(x.buf + off, x.ck + off, n, hipMemcpyDeviceToDevice, st));
        else HIP_TRY(hipMemcpyAsync(x.ck + off, x.buf + off, n, hipMemcpyDeviceToDevice, st));
    }
    return 0;
}

template <typename T, int FD>
static int forward_t(stfwi_plan *p, void *rec, int rec_mask, void *illum, int illum_comp, void *snap, int snap_mask,
                     hipStream_t st)
{
    FwdArgs<T> a = fwd_args<T>(p, rec, rec_mask, illum, illum_comp, snap, snap_mask);
    const int S = p->n_shots;
    fill<T>(p->state, (i64)S * 5 * p->fstride, st);
    fill<T>(p->psix_f, (i64)S * 5 * p->psix_len, st);
    fill<T>(p->psiz_f, (i64)S * 5 * p->psiz_len, st);
    const int rblk = rec ? (p->max_nrec + 255) / 256 : 0;
    int g0[STFWI_MAX_GROUPS + 1];
    const int G = group_bounds(p, g0);
    // with checkpoints the history holds the last segment only: frame 0 is
    // the carry-in of its first step
    const int K = p->ck;
    const int last0 = K ? ((p->nt_run - 1) / K) * K : 0;
    if (K && p->hist) {
        a.hist_t0 = last0 - 1;
        fill<T>(p->hist, (i64)S * p->hist_nt * 3 * p->hnp, st);
    }
    hipStream_t gs[STFWI_MAX_GROUPS];
    HIP_TRY(fork_groups(p, G, st, gs));
    const int rc = for_groups(p, G, st, [&](int k, GroupHandoff &h) -> int {
        FwdArgs<T> ak = a;
        ak.shot0 = g0[k];
        ak.nt_min = nt_min_of(p, g0[k], g0[k + 1]);
        const int ns = g0[k + 1] - g0[k];
        for (int it = 0; it < p->nt_run; ++it) {
            FwdArgs<T> b = ak;
            if (K && p->hist) {
                if (it > 0 && it % K == 0) {
                    const int rc2 = ckpt_copy<T>(p, it / K, g0[k], g0[k + 1], false, gs[k]);
                    if (rc2) return rc2;
                }
                if (it < a.hist_t0) {   // history of earlier segments is recomputed later
                    b.hist = nullptr;
                    b.hist2_mode = 0;
                }
            }
            if (it == 0) {
                h.wait_prev(k);
                HIP_TRY(stagger_in(p, k, gs));
            }
            launch_fwd<T, FD>(0, p->tz[0], b, rblk, ns, gs[k], it);
            if (it == 0) {
                HIP_TRY(stagger_out(p, k, G, gs));
                h.mark(k);
            }
            launch_fwd<T, FD>(1, p->tz[1], b, rblk, ns, gs[k], it);
        }
        // exx/ezz receivers of the last step
        if (rec) launch_fwd<T, FD>(0, p->tz[0], ak, rblk, ns, gs[k], p->nt_run);
        HIP_TRY(hipGetLastError());
        return 0;
    });
    if (rc) return rc;
    HIP_TRY(join_groups(p, G, st));
    return 0;
}

// zero `len` elements at ptr + k * pitch for k < rows (element size es)
static void fill_rows(void *ptr, i64 rows, i64 pitch, i64 len, int es, hipStream_t st)
{
    for (i64 k = 0; k < rows; ++k) {
        if (es == 8) fill<double>((double *)ptr + k * pitch, len, st);
        else fill<float>((float *)ptr + k * pitch, len, st);
    }
}

// Replay (capturing first when the key changed) of the work `body` enqueues
// on a stream; the replay is ordered after the caller's stream st and st
// after it.
template <typename F>
static int run_graph(stfwi_plan *p, int which, const std::vector<long long> &key, hipStream_t st, F body)
{
    if (!p->gexec[which] || p->gkey[which] != key) {
        if (p->gexec[which]) {
            // the previous executable may still be running on the capture
            // stream: let it finish before it is destroyed
            HIP_TRY(hipStreamSynchronize(p->cap));
            (void)hipGraphExecDestroy(p->gexec[which]);
            p->gexec[which] = nullptr;
        }
        HIP_TRY(hipStreamBeginCapture(p->cap, hipStreamCaptureModeThreadLocal));
        const int rc = body(p->cap);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(p->cap, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(STFWI_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
        const hipError_t e2 = hipGraphInstantiate(&p->gexec[which], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e2 != hipSuccess) {
            p->gexec[which] = nullptr;
            return fail(STFWI_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(e2));
        }
        p->gkey[which] = key;
    }
    HIP_TRY(hipEventRecord(p->ev_in, st));
    HIP_TRY(hipStreamWaitEvent(p->cap, p->ev_in, 0));
    HIP_TRY(hipGraphLaunch(p->gexec[which], p->cap));
    HIP_TRY(hipEventRecord(p->ev_out, p->cap));
    HIP_TRY(hipStreamWaitEvent(st, p->ev_out, 0));
    return 0;
}

static int forward_dispatch(stfwi_plan *p, void *rec, int rec_mask, void *illum, int illum_comp, void *snap,
                            int snap_mask, hipStream_t st)
{
    const int fd = p->cfg.fd_order;
    if (p->cfg.dtype == STFWI_F64)
        return fd == 8 ? forward_t<double, 8>(p, rec, rec_mask, illum, illum_comp, snap, snap_mask, st)
             : fd == 4 ? forward_t<double, 4>(p, rec, rec_mask, illum, illum_comp, snap, snap_mask, st)
                       : forward_t<double, 2>(p, rec, rec_mask, illum, illum_comp, snap, snap_mask, st);
    return fd == 8 ? forward_t<float, 8>(p, rec, rec_mask, illum, illum_comp, snap, snap_mask, st)
         : fd == 4 ? forward_t<float, 4>(p, rec, rec_mask, illum, illum_comp, snap, snap_mask, st)
                   : forward_t<float, 2>(p, rec, rec_mask, illum, illum_comp, snap, snap_mask, st);
}

extern "C" int stfwi_forward(stfwi_plan *p, void *rec, int rec_mask, void *illum, int illum_comp, void *snap,
                             int snap_mask, void *stream)
{
    if (!p) return fail(STFWI_ERR_INVALID, "null plan");
    if (!p->have_model) return fail(STFWI_ERR_STATE, "stfwi_set_model not called");
    if (!p->n_shots) return fail(STFWI_ERR_STATE, "stfwi_set_shots not called");
    if (illum && (illum_comp < 0 || illum_comp > 3)) return fail(STFWI_ERR_INVALID, "illum_comp=%d", illum_comp);
    hipStream_t st = (hipStream_t)stream;
    int rc;
    if (p->graphs && !illum && !snap) {
        // receivers go through rec_int: the requested channels are zeroed in
        // the graph and copied out after it
        const int mask = rec ? (rec_mask & 15) : 0;
        const i64 blk = (i64)p->cfg.nt_max * std::max(p->cfg.nrec_max, 1);
        const std::vector<long long> key = {p->n_shots, p->nt_run, p->tz[0], p->tz[1], p->groups, mask, p->max_nrec};
        rc = run_graph(p, 0, key, st, [&](hipStream_t cs) {
            for (int ch = 0; ch < 4; ++ch)
                if (mask & (1 << ch))
                    fill_rows((char *)p->rec_int + ch * blk * p->esz, p->n_shots, 4 * blk, blk, p->esz, cs);
            return forward_dispatch(p, mask ? p->rec_int : nullptr, mask, nullptr, 0, nullptr, 0, cs);
        });
        for (int ch = 0; rc == 0 && ch < 4; ++ch)
            if (mask & (1 << ch)) {
                const size_t row = (size_t)blk * p->esz, pitch = 4 * row;
                HIP_TRY(hipMemcpy2DAsync((char *)rec + ch * row, pitch, (char *)p->rec_int + ch * row, pitch, row,
                                         p->n_shots, hipMemcpyDeviceToDevice, st));
            }
    } else {
        rc = forward_dispatch(p, rec, rec_mask, illum, illum_comp, snap, snap_mask, st);
    }
    p->fwd_done = rc == 0 && p->hist != nullptr;
    return rc;
}

template <typename T, int FD>
static int adjoint_t(stfwi_plan *p, const void *rec_adj, void *src_adj, void *snap, int snap_mask, hipStream_t st)
{
    AdjArgs<T> a = adj_args<T>(p, rec_adj, src_adj, snap, snap_mask);
    const int S = p->n_shots;
    fill<T>(a.state, (i64)S * 5 * p->fstride, st);
    for (int k = 0; k < 2; ++k) {
        fill<T>(p->psix_a[k], (i64)S * 4 * p->psix_len, st);
        fill<T>(p->psiz_a[k], (i64)S * 4 * p->psiz_len, st);
    }
    if (a.imaging) fill<T>(p->gacc, (i64)S * p->nacc * p->hnp, st);
    // merged receiver adjoints of every step, read by the time-step kernels
    hipLaunchKernelGGL(k_merge_inj<T>, dim3((p->max_nuc + 127) / 128, std::min(p->nt_run, 65535), S), dim3(128), 0,
                       st, a, p->nt_run);
    int g0[STFWI_MAX_GROUPS + 1];
    const int G = group_bounds(p, g0);
    hipStream_t gs[STFWI_MAX_GROUPS];
    HIP_TRY(fork_groups(p, G, st, gs));
    // segments of the reverse sweep: one with the full history, or the
    // checkpoint segments last to first, each but the last recomputed
    const int K = (p->ck && a.imaging) ? p->ck : p->nt_run;
    const int nseg = (p->nt_run + K - 1) / K;
    FwdArgs<T> f = fwd_args<T>(p, nullptr, 0, nullptr, 0, nullptr, 0);
    f.dxu0 = nullptr;                 // kept from the first forward
    const int rc = for_groups(p, G, st, [&](int k, GroupHandoff &h) -> int {
        AdjArgs<T> ak = a;
        FwdArgs<T> fk = f;
        const int s0 = g0[k], ns = g0[k + 1] - g0[k];
        ak.shot0 = fk.shot0 = s0;
        ak.nt_min = fk.nt_min = nt_min_of(p, s0, g0[k + 1]);
        for (int seg = nseg - 1; seg >= 0; --seg) {
            const int t_lo = seg * K, t_hi = std::min(t_lo + K, p->nt_run);
            if (p->ck && a.imaging) {
                ak.hist_t0 = fk.hist_t0 = t_lo - 1;
                if (seg < nseg - 1) {
                    if (seg == 0) {
                        fill<T>((T *)p->state + (i64)s0 * 5 * p->fstride, (i64)ns * 5 * p->fstride, gs[k]);
                        fill<T>((T *)p->psix_f + (i64)s0 * 5 * p->psix_len, (i64)ns * 5 * p->psix_len, gs[k]);
                        fill<T>((T *)p->psiz_f + (i64)s0 * 5 * p->psiz_len, (i64)ns * 5 * p->psiz_len, gs[k]);
                    } else {
                        const int rc2 = ckpt_copy<T>(p, seg, s0, g0[k + 1], true, gs[k]);
                        if (rc2) return rc2;
                    }
                    const int nb = (int)std::min<i64>((p->np + 255) / 256, 1024);
                    hipLaunchKernelGGL(k_hist_from_state<T>, dim3(nb, ns), dim3(256), 0, gs[k], fk.g,
                                       (const T *)p->state, (i64)5 * p->fstride, (T *)p->hist,
                                       (const ShotDev *)p->shots_d, 0, s0);
                    for (int it = t_lo; it < t_hi; ++it) {
                        launch_fwd<T, FD>(0, p->tz[0], fk, 0, ns, gs[k], it);
                        launch_fwd<T, FD>(1, p->tz[1], fk, 0, ns, gs[k], it);
                    }
                }
            }
            for (int it = t_hi - 1; it >= t_lo; --it) {
                const int cur = it & 1, prev = cur ^ 1;
                ak.psix_o = (const T *)p->psix_a[prev]; ak.psiz_o = (const T *)p->psiz_a[prev];
                ak.psix_n = (T *)p->psix_a[cur]; ak.psiz_n = (T *)p->psiz_a[cur];
                bool first = false;
                for (int s = s0; s < g0[k + 1]; ++s) first |= p->shot_nt[s] - 1 == it;
                if (first)
                    hipLaunchKernelGGL(k_adj_inject<T>, dim3((p->max_nuc + 63) / 64, 1, ns), dim3(64), 0, gs[k], ak,
                                       it);
                const bool start = seg == nseg - 1 && it == t_hi - 1;
                if (start) {
                    h.wait_prev(k);
                    HIP_TRY(stagger_in(p, k, gs));
                }
                launch_adj<T, FD>(2, p->tz[2], ak, ns, gs[k], it);
                if (start) {
                    HIP_TRY(stagger_out(p, k, G, gs));
                    h.mark(k);
                }
                launch_adj<T, FD>(3, p->tz[3], ak, ns, gs[k], it);
            }
        }
        HIP_TRY(hipGetLastError());
        return 0;
    });
    if (rc) return rc;
    HIP_TRY(join_groups(p, G, st));
    return 0;
}

extern "C" int stfwi_adjoint(stfwi_plan *p, const void *rec_adj, void *src_adj, void *snap, int snap_mask,
                             void *stream)
{
    if (!p || !rec_adj) return fail(STFWI_ERR_INVALID, "null argument");
    if (!p->have_model || !p->n_shots) return fail(STFWI_ERR_STATE, "model or shots not set");
    if (p->cfg.imaging && p->hist && !p->fwd_done)
        return fail(STFWI_ERR_STATE, "imaging needs a stfwi_forward with the same shots first");
    hipStream_t st = (hipStream_t)stream;
    const int fd = p->cfg.fd_order;
    auto dispatch = [&](const void *ra, void *sa, void *sn, int sm, hipStream_t s2) {
        if (p->cfg.dtype == STFWI_F64)
            return fd == 8 ? adjoint_t<double, 8>(p, ra, sa, sn, sm, s2)
                 : fd == 4 ? adjoint_t<double, 4>(p, ra, sa, sn, sm, s2)
                           : adjoint_t<double, 2>(p, ra, sa, sn, sm, s2);
        return fd == 8 ? adjoint_t<float, 8>(p, ra, sa, sn, sm, s2)
             : fd == 4 ? adjoint_t<float, 4>(p, ra, sa, sn, sm, s2)
                       : adjoint_t<float, 2>(p, ra, sa, sn, sm, s2);
    };
    int rc;
    if (p->graphs && !snap && !p->aen) {
        const size_t rbytes = (size_t)p->n_shots * 4 * p->cfg.nt_max * std::max(p->cfg.nrec_max, 1) * p->esz;
        HIP_TRY(hipMemcpyAsync(p->radj_int, rec_adj, rbytes, hipMemcpyDeviceToDevice, st));
        const int want = src_adj ? 1 : 0;
        // the launch sequence depends on every shot's record length (the
        // separate first-step injection), so all of them are in the key
        std::vector<long long> key = {p->n_shots, p->nt_run, p->tz[2], p->tz[3], p->groups, want, p->max_nuc};
        key.insert(key.end(), p->shot_nt.begin(), p->shot_nt.end());
        rc = run_graph(p, 1, key, st, [&](hipStream_t cs) {
            if (want) fill_rows(p->sadj_int, 1, 0, (i64)p->n_shots * p->cfg.nt_max, p->esz, cs);
            return dispatch(p->radj_int, want ? p->sadj_int : nullptr, nullptr, 0, cs);
        });
        if (rc == 0 && want)
            HIP_TRY(hipMemcpyAsync(src_adj, p->sadj_int, (size_t)p->n_shots * p->cfg.nt_max * p->esz,
                                   hipMemcpyDeviceToDevice, st));
    } else {
        rc = dispatch(rec_adj, src_adj, snap, snap_mask, st);
    }
    p->adj_done = rc == 0 && p->gacc != nullptr && p->cfg.imaging != 0;
    return rc;
}

template <typename T, typename C>
static int finalize_t(stfwi_plan *p, const void *vs, const void *vp, const void *rho, void *out, hipStream_t st)
{
    const int blocks = (int)((p->np + 255) / 256);
    hipLaunchKernelGGL((k_finalize<T, C>), dim3(blocks), dim3(256), 0, st, p->n_shots, p->np, p->nacc,
                       p->cfg.imaging, p->cfg.grad_params, (T)p->cfg.dt, (const T *)p->gacc, (const C *)vs,
                       (const C *)vp, (const C *)rho, (T *)out, p->cfg.nx_dom, p->hpitch, p->hcoff, p->hnp);
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int stfwi_gradient_finalize(stfwi_plan *p, const void *vs, const void *vp, const void *rho,
                                       int chain_dtype, void *g_out, void *stream)
{
    if (!p || !vs || !vp || !rho || !g_out) return fail(STFWI_ERR_INVALID, "null argument");
    if (!p->adj_done) return fail(STFWI_ERR_STATE, "no imaging adjoint has run");
    hipStream_t st = (hipStream_t)stream;
    const bool t64 = p->cfg.dtype == STFWI_F64;
    if (chain_dtype == STFWI_F64)
        return t64 ? finalize_t<double, double>(p, vs, vp, rho, g_out, st) : finalize_t<float, double>(p, vs, vp, rho, g_out, st);
    if (chain_dtype == STFWI_F32)
        return t64 ? finalize_t<double, float>(p, vs, vp, rho, g_out, st) : finalize_t<float, float>(p, vs, vp, rho, g_out, st);
    return fail(STFWI_ERR_INVALID, "chain_dtype=%d", chain_dtype);
}

template <typename T>
static int residual_t(stfwi_plan *p, const void *pred, const void *obs, int comp, const double *sos, int n_sos,
                      int n_taper, double scale, int n_norm, void *loss, void *rec_adj, hipStream_t st)
{
    ResArgs<T> a;
    a.shots = p->shots_d;
    a.pred = (const T *)pred; a.obs = (const T *)obs; a.rec_adj = (T *)rec_adj; a.loss = (T *)loss;
    a.nt_max = p->cfg.nt_max; a.nrec_max = p->cfg.nrec_max; a.comp = comp;
    a.n_sos = n_sos; a.n_taper = n_taper;
    for (int k = 0; k < n_sos; ++k) {
        const double *s = sos + 6 * k;
        a.coef[k][0] = (T)(s[0] / s[3]); a.coef[k][1] = (T)(s[1] / s[3]); a.coef[k][2] = (T)(s[2] / s[3]);
        a.coef[k][3] = (T)(s[4] / s[3]); a.coef[k][4] = (T)(s[5] / s[3]);
    }
    a.w = (T)(scale / n_norm);
    const i64 chs = (i64)a.nt_max * a.nrec_max, n = (i64)p->n_shots * 4 * chs;
    hipLaunchKernelGGL(k_residual_zero<T>, dim3((int)std::min<i64>((n + 255) / 256, 16384)), dim3(256), 0, st,
                       (T *)rec_adj, n, chs, comp);
    int nch = 1;
    if (n_sos == 0) {
        // no filter: parallel over blocks of time steps as well
        nch = (a.nt_max + RES_TCH - 1) / RES_TCH;
        hipLaunchKernelGGL(k_residual_el<T>, dim3((a.nrec_max + 63) / 64, nch, p->n_shots), dim3(64), 0, st, a,
                           p->lpart);
    } else {
        hipLaunchKernelGGL(k_residual<T>, dim3((a.nrec_max + 63) / 64, p->n_shots), dim3(64), 0, st, a, p->lpart);
    }
    hipLaunchKernelGGL(k_residual_loss<T>, dim3(p->n_shots), dim3(256), 0, st, a, (const double *)p->lpart, nch);
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int stfwi_residual(stfwi_plan *p, const void *pred, const void *obs, int component, const double *sos,
                              int n_sos, int n_taper, double scale, int n_shots_norm, void *loss_out, void *rec_adj,
                              void *stream)
{
    if (!p || !pred || !obs || !loss_out || !rec_adj) return fail(STFWI_ERR_INVALID, "null argument");
    if (!p->n_shots) return fail(STFWI_ERR_STATE, "stfwi_set_shots not called");
    if (component < 0 || component > 3) return fail(STFWI_ERR_INVALID, "component=%d", component);
    if (n_sos < 0 || n_sos > MAX_SOS || (n_sos > 0 && !sos)) return fail(STFWI_ERR_INVALID, "n_sos=%d", n_sos);
    if (n_shots_norm < 1) return fail(STFWI_ERR_INVALID, "n_shots_norm=%d", n_shots_norm);
    for (int k = 0; k < n_sos; ++k)
        if (sos[6 * k + 3] == 0.0) return fail(STFWI_ERR_INVALID, "sos section %d has a0 = 0", k);
    hipStream_t st = (hipStream_t)stream;
    return p->cfg.dtype == STFWI_F64
               ? residual_t<double>(p, pred, obs, component, sos, n_sos, n_taper, scale, n_shots_norm, loss_out, rec_adj, st)
               : residual_t<float>(p, pred, obs, component, sos, n_sos, n_taper, scale, n_shots_norm, loss_out, rec_adj, st);
}

// ---------------------------------------------------------------------------
// diagnostics: time `reps` launches of one time-step kernel for all shots of
// the plan at a mid-record step, bracketed by HIP events on `stream`.  With
// grouped = true each repetition launches every shot group on its own
// stream, as the time loops do (the autotuner times candidates this way);
// otherwise one launch covers all shots.
// ---------------------------------------------------------------------------
template <typename T, int FD>
static int time_kernel_t(stfwi_plan *p, int which, int tz, int reps, hipStream_t st, double *ms, bool grouped)
{
    const int it = p->nt_run / 2;
    int g0[STFWI_MAX_GROUPS + 1] = {0, p->n_shots};
    const int G = grouped ? group_bounds(p, g0) : 1;
    hipStream_t gs[STFWI_MAX_GROUPS];
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    if (which <= 1) {
        FwdArgs<T> a = fwd_args<T>(p, nullptr, 0, nullptr, 0, nullptr, 0);
        if (p->ck) a.hist_t0 = it - 1;    // frames 0..2 of a checkpoint segment
        HIP_TRY(hipEventRecord(e0, st));
        HIP_TRY(fork_groups(p, G, st, gs));
        for (int r = 0; r < reps; ++r)
            for (int k = 0; k < G; ++k) {
                a.shot0 = g0[k];
                a.nt_min = nt_min_of(p, g0[k], g0[k + 1]);
                launch_fwd<T, FD>(which, tz, a, 0, g0[k + 1] - g0[k], gs[k], it);
            }
        HIP_TRY(join_groups(p, G, st));
        HIP_TRY(hipEventRecord(e1, st));
    } else {
        AdjArgs<T> a = adj_args<T>(p, nullptr, nullptr, nullptr, 0);
        if (p->ck) a.hist_t0 = it - 1;
        a.psix_o = (const T *)p->psix_a[0]; a.psiz_o = (const T *)p->psiz_a[0];
        a.psix_n = (T *)p->psix_a[1]; a.psiz_n = (T *)p->psiz_a[1];
        HIP_TRY(hipEventRecord(e0, st));
        HIP_TRY(fork_groups(p, G, st, gs));
        for (int r = 0; r < reps; ++r)
            for (int k = 0; k < G; ++k) {
                a.shot0 = g0[k];
                a.nt_min = nt_min_of(p, g0[k], g0[k + 1]);
                launch_adj<T, FD>(which, tz, a, g0[k + 1] - g0[k], gs[k], it);
            }
        HIP_TRY(join_groups(p, G, st));
        HIP_TRY(hipEventRecord(e1, st));
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventSynchronize(e1));
    float f = 0;
    HIP_TRY(hipEventElapsedTime(&f, e0, e1));
    *ms = f;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 0;
}

static int time_kernel(stfwi_plan *p, int which, int tz, int reps, hipStream_t st, double *ms,
                       bool grouped = false)
{
    const int fd = p->cfg.fd_order;
    if (p->cfg.dtype == STFWI_F64)
        return fd == 8 ? time_kernel_t<double, 8>(p, which, tz, reps, st, ms, grouped)
             : fd == 4 ? time_kernel_t<double, 4>(p, which, tz, reps, st, ms, grouped)
                       : time_kernel_t<double, 2>(p, which, tz, reps, st, ms, grouped);
    return fd == 8 ? time_kernel_t<float, 8>(p, which, tz, reps, st, ms, grouped)
         : fd == 4 ? time_kernel_t<float, 4>(p, which, tz, reps, st, ms, grouped)
                   : time_kernel_t<float, 2>(p, which, tz, reps, st, ms, grouped);
}

extern "C" int stfwi_time_kernel(stfwi_plan *p, int which, int reps, int grouped, void *stream, double *ms)
{
    if (!p || !ms || which < 0 || which > 3 || reps < 1) return fail(STFWI_ERR_INVALID, "bad argument");
    if (!p->have_model || !p->n_shots) return fail(STFWI_ERR_STATE, "model or shots not set");
    const int rc = time_kernel(p, which, p->tz[which], reps, (hipStream_t)stream, ms, grouped != 0);
    // the timing launches ran on the live buffers: a pending forward or
    // adjoint is no longer valid
    p->fwd_done = false;
    p->adj_done = false;
    return rc;
}

extern "C" int stfwi_autotune(stfwi_plan *p, int reps, void *stream, int *tz_out, double *us_out)
{
    if (!p || reps < 1) return fail(STFWI_ERR_INVALID, "bad argument");
    if (!p->have_model || !p->n_shots) return fail(STFWI_ERR_STATE, "model or shots not set");
    const int cand[4] = {4, 8, 16, 32};
    const bool f64 = p->cfg.dtype == STFWI_F64;
    for (int which = 0; which < 4; ++which) {
        // three interleaved rounds, the fastest round of each candidate: a
        // transient slowdown during one round does not decide the choice.
        // Each candidate runs as the time loops run it: the plan's shot
        // groups launched concurrently on their streams.
        double t[4] = {1e300, 1e300, 1e300, 1e300};
        for (int round = 0; round < 3; ++round)
            for (int ci = 0; ci < 4; ++ci) {
                const int tz = cand[ci];
                if (f64 ? !tz_supported<double>(tz) : !tz_supported<float>(tz)) continue;
                double ms = 0;
                int rc = time_kernel(p, which, tz, 2, (hipStream_t)stream, &ms, true);   // warm-up
                if (!rc) rc = time_kernel(p, which, tz, reps, (hipStream_t)stream, &ms, true);
                if (rc) return rc;
                t[ci] = std::min(t[ci], ms);
            }
        double best = 1e300;
        int best_tz = p->tz[which];
        for (int ci = 0; ci < 4; ++ci) {
            if (t[ci] >= 1e300) continue;
            if (us_out) us_out[which * 4 + ci] = t[ci] * 1e3 / reps;
            if (t[ci] < best) {
                best = t[ci];
                best_tz = cand[ci];
            }
        }
        p->tz[which] = best_tz;
        if (tz_out) tz_out[which] = best_tz;
    }
    // the timing launches ran on the live buffers
    p->fwd_done = false;
    p->adj_done = false;
    return 0;
}

extern "C" int stfwi_set_adjoint_energy(stfwi_plan *p, void *energy, int comp)
{
    if (!p || comp < 0 || comp > 4) return fail(STFWI_ERR_INVALID, "bad argument (comp=%d)", comp);
    p->aen = energy;
    p->aen_comp = comp;
    return 0;
}

extern "C" int stfwi_set_lean(stfwi_plan *p, int enable)
{
    if (!p) return fail(STFWI_ERR_INVALID, "null plan");
    p->lean = enable ? 1 : 0;
    return 0;
}

extern "C" int stfwi_set_groups(stfwi_plan *p, int groups)
{
    if (!p || groups < 1 || groups > STFWI_MAX_GROUPS)
        return fail(STFWI_ERR_INVALID, "groups=%d outside [1, %d]", groups, STFWI_MAX_GROUPS);
    p->groups = groups;
    return 0;
}

extern "C" int stfwi_set_tiles(stfwi_plan *p, const int *tz)
{
    if (!p || !tz) return fail(STFWI_ERR_INVALID, "bad argument");
    for (int k = 0; k < 4; ++k) {
        const bool ok = p->cfg.dtype == STFWI_F64 ? tz_supported<double>(tz[k]) : tz_supported<float>(tz[k]);
        if (!ok) return fail(STFWI_ERR_UNSUPPORTED, "tile height %d not compiled for this precision", tz[k]);
    }
    for (int k = 0; k < 4; ++k) p->tz[k] = tz[k];
    return 0;
}

// ---------------------------------------------------------------------------
// plan-free post-processing entry points
// ---------------------------------------------------------------------------
template <typename T>
static int smooth_t(const void *in, void *tmp, void *out, int nz, int nx, double sigma, hipStream_t st)
{
    const int r = (int)(3.0 * sigma + 0.5);
    if (r > MAX_SMOOTH_R) return fail(STFWI_ERR_INVALID, "sigma=%g: kernel radius %d > %d", sigma, r, MAX_SMOOTH_R);
    SmoothW w;
    double sum = 0;
    for (int k = -r; k <= r; ++k) {
        const double x = k / sigma;
        w.w[k + r] = exp(-0.5 * x * x);
        sum += w.w[k + r];
    }
    for (int k = 0; k <= 2 * r; ++k) w.w[k] /= sum;
    dim3 grid((nx + 127) / 128, nz);
    hipLaunchKernelGGL(k_smooth_rows<T>, grid, dim3(128), 0, st, (const T *)in, (T *)tmp, nz, nx, r, w);
    hipLaunchKernelGGL(k_smooth_cols<T>, grid, dim3(128), 0, st, (const T *)tmp, (T *)out, nz, nx, r, w);
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int stfwi_smooth2d(const void *in, void *tmp, void *out, int nz, int nx, double sigma, int dtype,
                              void *stream)
{
    if (!in || !tmp || !out || nz < 1 || nx < 1 || !(sigma > 0)) return fail(STFWI_ERR_INVALID, "bad argument");
    if (tmp == in || tmp == out) return fail(STFWI_ERR_INVALID, "tmp must not alias in/out");
    return dtype == STFWI_F64 ? smooth_t<double>(in, tmp, out, nz, nx, sigma, (hipStream_t)stream)
                              : smooth_t<float>(in, tmp, out, nz, nx, sigma, (hipStream_t)stream);
}

extern "C" int stfwi_illumination(const void *ws, int n_shots, int nz_t, int nx_t, int z0, int pad, int nz_c,
                                  int nx_c, double eps, int dtype, void *out, void *stream)
{
    if (!ws || !out || n_shots < 1 || nz_c < 1 || nx_c < 1 || z0 + nz_c > nz_t || pad + nx_c > nx_t)
        return fail(STFWI_ERR_INVALID, "bad argument");
    hipStream_t st = (hipStream_t)stream;
    if (dtype == STFWI_F64)
        hipLaunchKernelGGL(k_illum_hinv<double>, dim3(1), dim3(256), 0, st, (const double *)ws, n_shots, nz_t, nx_t,
                           z0, pad, nz_c, nx_c, eps, (double *)out);
    else
        hipLaunchKernelGGL(k_illum_hinv<float>, dim3(1), dim3(256), 0, st, (const float *)ws, n_shots, nz_t, nx_t,
                           z0, pad, nz_c, nx_c, eps, (float *)out);
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int stfwi_eprecond3(const void *ws, int n_shots, int nz_t, int nx_t, int free_surface, int pad, int nz_dom,
                               int nx_dom, const int *rec_x, const int *rec_z, int nrec, double dx, double dz, double eps,
                               int dtype, void *we_pad, void *out, void *stream)
{
    if (!ws || !we_pad || !out || !rec_x || !rec_z || nrec < 1 || n_shots < 1)
        return fail(STFWI_ERR_INVALID, "bad argument");
    // orientation of the receiver line: the coordinate with the larger spread
    double mx = 0, mz = 0;
    for (int i = 0; i < nrec; ++i) { mx += rec_x[i]; mz += rec_z[i]; }
    mx /= nrec; mz /= nrec;
    double vx = 0, vz = 0;
    for (int i = 0; i < nrec; ++i) {
        vx += (rec_x[i] - mx) * (rec_x[i] - mx);
        vz += (rec_z[i] - mz) * (rec_z[i] - mz);
    }
    const int horiz = vx >= vz;
    const int *ln = horiz ? rec_x : rec_z;
    int lo = ln[0], hi = ln[0];
    for (int i = 1; i < nrec; ++i) { lo = std::min(lo, ln[i]); hi = std::max(hi, ln[i]); }
    const double s_line = horiz ? dx : dz, s_perp = horiz ? dz : dx;
    const double lmin = lo * s_line, lmax = hi * s_line, perp = (horiz ? rec_z[0] : rec_x[0]) * s_perp;
    const int z0 = free_surface ? 0 : pad;
    hipStream_t st = (hipStream_t)stream;
    if (dtype == STFWI_F64)
        hipLaunchKernelGGL(k_eprecond3<double>, dim3(1), dim3(256), 0, st, (const double *)ws, n_shots, nz_t, nx_t, z0,
                           pad, nz_dom, nx_dom, horiz, lmin, lmax, perp, dx, dz, eps, (double *)we_pad, (double *)out);
    else
        hipLaunchKernelGGL(k_eprecond3<float>, dim3(1), dim3(256), 0, st, (const float *)ws, n_shots, nz_t, nx_t, z0,
                           pad, nz_dom, nx_dom, horiz, lmin, lmax, perp, dx, dz, eps, (float *)we_pad, (float *)out);
    HIP_TRY(hipGetLastError());
    return 0;
}

// blocks of 256 threads for a grid-stride loop over n elements
static int grid1(long long n) { return (int)std::min<long long>((n + 255) / 256, 8192); }

// blocks of the first pass of a min/max reduction: at most MINMAX_PARTS
// partials per call
#define MINMAX_PARTS 1024

// min (op = 0) or max (op = 1) of x[0, n) into *res: two passes through
// `part` (MINMAX_PARTS elements)
template <typename T>
static int dev_minmax_into(const T *x, long long n, int op, T *res, T *part, hipStream_t st)
{
    const int nb = (int)std::min<long long>((n + 255) / 256, MINMAX_PARTS);
    hipLaunchKernelGGL(k_minmax_part<T>, dim3(nb), dim3(256), 0, st, x, n, op, part);
    hipLaunchKernelGGL(k_minmax_part<T>, dim3(1), dim3(256), 0, st, (const T *)part, (long long)nb, op, res);
    HIP_TRY(hipGetLastError());
    return 0;
}

// eprecond1 with its reduction scratch given: `part` holds MINMAX_PARTS
// partial maxima and `mx` the maximum
template <typename T>
static int eprecond1_core(const T *ws, const T *wr, long long n, double eps, T *we, T *part, T *mx, hipStream_t st)
{
    hipLaunchKernelGGL(k_sqrt_prod<T>, dim3(grid1(n)), dim3(256), 0, st, ws, wr, n, we);
    const int rc = dev_minmax_into<T>(we, n, 1, mx, part, st);
    if (rc == 0) hipLaunchKernelGGL(k_add_eps_max<T>, dim3(grid1(n)), dim3(256), 0, st, we, n, (const T *)mx, eps);
    HIP_TRY(hipGetLastError());
    return rc;
}

// the scratch from the stream-ordered pool, freed after the kernels
template <typename T>
static int eprecond1_t(const T *ws, const T *wr, long long n, double eps, T *we, hipStream_t st)
{
    T *buf = nullptr;
    HIP_TRY(hipMallocAsync((void **)&buf, (MINMAX_PARTS + 1) * sizeof(T), st));
    const int rc = eprecond1_core<T>(ws, wr, n, eps, we, buf, buf + MINMAX_PARTS, st);
    HIP_TRY(hipFreeAsync(buf, st));
    return rc;
}

extern "C" int stfwi_eprecond1(const void *ws, const void *wr, long long n, double eps, int dtype, void *we, void *stream)
{
    if (!ws || !wr || !we || n < 1) return fail(STFWI_ERR_INVALID, "bad argument");
    hipStream_t st = (hipStream_t)stream;
    if (dtype == STFWI_F64) return eprecond1_t<double>((const double *)ws, (const double *)wr, n, eps, (double *)we, st);
    if (dtype == STFWI_F32) return eprecond1_t<float>((const float *)ws, (const float *)wr, n, eps, (float *)we, st);
    return fail(STFWI_ERR_INVALID, "dtype=%d", dtype);
}

// ---------------------------------------------------------------------------
// allocation-free eprecond1: the reduction scratch is a file-scope device
// array, one copy per GPU, held until the process ends.  Its host-side name
// is not a device address: host code uses only the per-device address from
// hipGetSymbolAddress.  Every call on a GPU shares that GPU's copy, so calls
// that overlap on one GPU (other streams, other threads) race.  Nothing in
// it needs to start at zero: the first reduction pass writes every partial
// the second reads, and the second writes the maximum.
// ---------------------------------------------------------------------------
#define NOALLOC_SCRATCH_DOUBLES (MINMAX_PARTS + 1)
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

#define STFWI_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[STFWI_MAX_DEVICES] = {};

// the current device's copy of `scratch`, looked up on the first call there
static int scratch_for_current_device(double **out)
{
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= STFWI_MAX_DEVICES) return fail(STFWI_ERR_INVALID, "device %d outside [0, %d)", dev,
                                                         STFWI_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        double *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(STFWI_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", dev,
                        hipGetErrorString(e));
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

extern "C" int stfwi_eprecond1_noalloc(const void *ws, const void *wr, long long n, double eps, int dtype, void *we,
                                       void *stream)
{
    if (!ws || !wr || !we || n < 1) return fail(STFWI_ERR_INVALID, "bad argument");
    if (dtype != STFWI_F64 && dtype != STFWI_F32) return fail(STFWI_ERR_INVALID, "dtype=%d", dtype);
    // the reduction needs min(ceil(n / 256), MINMAX_PARTS) partials plus the
    // maximum: never more than the scratch holds, for any n
    const long long need = std::min<long long>((n + 255) / 256, MINMAX_PARTS) + 1;
    if (need > NOALLOC_SCRATCH_DOUBLES)
        return fail(STFWI_ERR_INVALID, "n=%lld needs %lld scratch elements, more than the %d held", n, need,
                    NOALLOC_SCRATCH_DOUBLES);
    double *ptr = nullptr;
    const int rc = scratch_for_current_device(&ptr);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    if (dtype == STFWI_F64)
        return eprecond1_core<double>((const double *)ws, (const double *)wr, n, eps, (double *)we, ptr,
                                      ptr + MINMAX_PARTS, st);
    float *fp = (float *)ptr;
    return eprecond1_core<float>((const float *)ws, (const float *)wr, n, eps, (float *)we, fp, fp + MINMAX_PARTS, st);
}

extern "C" int stfwi_taper_grad(int nz, int nx, double dz, int gradt1, int gradt2, double exp_taper, int bottom,
                                int gradb1, int gradb2, double *out, void *stream)
{
    if (!out || nz < 1 || nx < 1) return fail(STFWI_ERR_INVALID, "bad argument");
    const long long n = (long long)nz * nx;
    hipLaunchKernelGGL(k_taper_grad, dim3(grid1(n)), dim3(256), 0, (hipStream_t)stream, nz, nx, dz, gradt1, gradt2,
                       exp_taper, bottom, gradb1, gradb2, out);
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int stfwi_source_taper(int nz, int nx, double dh, const int *src_xz, int nsrc, double srt_radius,
                                  int filt_size, double *out, void *stream)
{
    if (!out || nz < 1 || nx < 1 || nsrc < 0 || (nsrc > 0 && !src_xz) || !(dh > 0) || !(srt_radius > 0))
        return fail(STFWI_ERR_INVALID, "bad argument");
    const int ps = std::max((int)std::ceil(2.0 * srt_radius / dh), 5);
    if ((long long)ps * ps > (1LL << 26)) return fail(STFWI_ERR_INVALID, "taper radius too large for the spacing");
    hipStream_t st = (hipStream_t)stream;
    const long long n = (long long)nz * nx;
    double *patch = nullptr, *mm = nullptr;
    int *src = nullptr;
    HIP_TRY(hipMallocAsync((void **)&patch, (size_t)ps * ps * sizeof(double), st));
    HIP_TRY(hipMallocAsync((void **)&mm, (2 + MINMAX_PARTS) * sizeof(double), st));
    HIP_TRY(hipMallocAsync((void **)&src, (size_t)std::max(nsrc, 1) * 2 * sizeof(int), st));
    if (nsrc > 0) HIP_TRY(hipMemcpyAsync(src, src_xz, (size_t)nsrc * 2 * sizeof(int), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_src_patch, dim3(1), dim3(256), 0, st, ps, dh, srt_radius, patch);
    hipLaunchKernelGGL(k_src_stamp, dim3(grid1(n)), dim3(256), 0, st, nz, nx, ps, (const double *)patch,
                       (const int *)src, nsrc, out);
    int rc = dev_minmax_into<double>(out, n, 0, mm, mm + 2, st);
    if (rc == 0) rc = dev_minmax_into<double>(out, n, 1, mm + 1, mm + 2, st);
    if (rc == 0)
        hipLaunchKernelGGL(k_src_finish, dim3(grid1(n)), dim3(256), 0, st, nz, nx, (const double *)mm,
                           (const int *)src, nsrc, filt_size, out);
    HIP_TRY(hipFreeAsync(patch, st));
    HIP_TRY(hipFreeAsync(mm, st));
    HIP_TRY(hipFreeAsync(src, st));
    HIP_TRY(hipGetLastError());
    // the host source list is read by the copy above before this returns
    HIP_TRY(hipStreamSynchronize(st));
    return rc;
}

template <typename T>
static int sos_t(void *x, int nt, int ntr, const double *sos, int n_sos, int n_taper, int transpose, hipStream_t st)
{
    SosArgs<T> a;
    a.x = (T *)x; a.nt = nt; a.ntr = ntr; a.n_sos = n_sos; a.n_taper = n_taper; a.transpose = transpose;
    for (int k = 0; k < n_sos; ++k) {
        const double *s = sos + 6 * k;
        a.coef[k][0] = (T)(s[0] / s[3]); a.coef[k][1] = (T)(s[1] / s[3]); a.coef[k][2] = (T)(s[2] / s[3]);
        a.coef[k][3] = (T)(s[4] / s[3]); a.coef[k][4] = (T)(s[5] / s[3]);
    }
    hipLaunchKernelGGL(k_sos_traces<T>, dim3((ntr + 127) / 128), dim3(128), 0, st, a);
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int stfwi_sosfilt(void *x, int nt, int ntr, const double *sos, int n_sos, int n_taper, int transpose,
                             int dtype, void *stream)
{
    if (!x || nt < 1 || ntr < 1 || n_sos < 0 || n_sos > 8 || (n_sos && !sos)) return fail(STFWI_ERR_INVALID, "bad argument");
    for (int k = 0; k < n_sos; ++k)
        if (sos[6 * k + 3] == 0.0) return fail(STFWI_ERR_INVALID, "sos section %d has a0 = 0", k);
    return dtype == STFWI_F64 ? sos_t<double>(x, nt, ntr, sos, n_sos, n_taper, transpose, (hipStream_t)stream)
                              : sos_t<float>(x, nt, ntr, sos, n_sos, n_taper, transpose, (hipStream_t)stream);
}

extern "C" int stfwi_zhou_imaging(const void *ex, const void *ez, const void *es, const void *ua, const void *va,
                                  const void *b_full, const void *vx, const void *vz, int nt, int nz, int nx,
                                  double dx, double dz, double dt, int dtype, void *out, void *stream)
{
    if (!ex || !ez || !es || !ua || !va || !b_full || !out || nt < 1 || nz < 1 || nx < 1 || (!vx) != (!vz))
        return fail(STFWI_ERR_INVALID, "bad argument");
    hipStream_t st = (hipStream_t)stream;
    dim3 grid((nx + 127) / 128, nz);
    if (dtype == STFWI_F64)
        hipLaunchKernelGGL(k_zhou_cubes<double>, grid, dim3(128), 0, st, (const double *)ex, (const double *)ez,
                           (const double *)es, (const double *)ua, (const double *)va, (const double *)b_full,
                           (const double *)vx, (const double *)vz, nt, nz, nx, dx, dz, dt, (double *)out);
    else
        hipLaunchKernelGGL(k_zhou_cubes<float>, grid, dim3(128), 0, st, (const float *)ex, (const float *)ez,
                           (const float *)es, (const float *)ua, (const float *)va, (const float *)b_full,
                           (const float *)vx, (const float *)vz, nt, nz, nx, (float)dx, (float)dz, (float)dt,
                           (float *)out);
    HIP_TRY(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// L-BFGS (float64)
// ---------------------------------------------------------------------------
extern "C" int stfwi_lbfgs_direction(const double *g, const double *S, const double *Y, const double *rho,
                                     const int *slots, int m, long long n, long long ld, double gamma, double *d,
                                     double *gd, void *stream)
{
    if (!g || !d || n < 1 || m < 0 || m > LB_MAX_M || (m > 0 && (!S || !Y || !rho || !slots || ld < n)))
        return fail(STFWI_ERR_INVALID, "bad argument (m=%d, n=%lld, ld=%lld)", m, n, ld);
    if (!(gamma > 0)) return fail(STFWI_ERR_INVALID, "gamma must be > 0");
    LbHist h;
    memset(&h, 0, sizeof h);
    for (int k = 0; k < m; ++k) {
        if (slots[k] < 0) return fail(STFWI_ERR_INVALID, "negative history slot");
        h.slot[k] = slots[k];
        h.rho[k] = rho[k];
    }
    hipLaunchKernelGGL(k_lbfgs_dir, dim3(1), dim3(LB_THREADS), 0, (hipStream_t)stream, g, S, Y, h, m, n, ld, gamma,
                       d, gd);
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int stfwi_lbfgs_pair(const double *x_new, const double *x_old, const double *g_new, const double *g_old,
                                long long n, double *s_out, double *y_out, double *dots, void *stream)
{
    if (!x_new || !x_old || !g_new || !g_old || !s_out || !y_out || !dots || n < 1)
        return fail(STFWI_ERR_INVALID, "bad argument");
    hipLaunchKernelGGL(k_lbfgs_pair, dim3(1), dim3(LB_THREADS), 0, (hipStream_t)stream, x_new, x_old, g_new, g_old,
                       n, s_out, y_out, dots);
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int stfwi_dot64(const double *a, const double *b, long long n, double *out, void *stream)
{
    if (!a || !b || !out || n < 1) return fail(STFWI_ERR_INVALID, "bad argument");
    hipLaunchKernelGGL(k_dot64, dim3(1), dim3(LB_THREADS), 0, (hipStream_t)stream, a, b, n, out);
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int stfwi_step_clip(const double *x, const double *d, double t, const double *lo, const double *hi,
                               long long n, double *out, void *stream)
{
    if (!x || !d || !out || n < 1) return fail(STFWI_ERR_INVALID, "bad argument");
    const long long blocks = std::min<long long>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_step_clip, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, x, d, t, lo, hi, n,
                       out);
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int stfwi_bound_mask(const double *x, const double *g, const double *lo, const double *hi, long long n,
                                double *v, void *stream)
{
    if (!x || !g || !v || n < 1) return fail(STFWI_ERR_INVALID, "bad argument");
    const long long blocks = std::min<long long>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_bound_mask, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, x, g, lo, hi, n, v);
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int stfwi_set_graphs(stfwi_plan *p, int enable)
{
    if (!p) return fail(STFWI_ERR_INVALID, "null plan");
    // Graph replay is refused: with it on, a GPU test run ended in a host
    // segmentation fault inside stfwi_adjoint after the capture key changed,
    // and no mistake in this library has been shown to cause it.  The eager
    // time loops are the only path until one is.
    if (enable) return fail(STFWI_ERR_UNSUPPORTED, "graph replay of the time loops is disabled");
    p->graphs = false;
    return 0;
}

// ---------------------------------------------------------------------------
// learned preconditioner: BF16-MFMA convolutions and FP32 layer kernels
// ---------------------------------------------------------------------------

// the LDS-tiled convolution: the input channels of the call run in passes
// of at most `chunk` channels (the first pass writes the output, the others
// add to it, the last adds the bias and applies the ReLU).  Each pass packs
// its slice of the weights into a stream-ordered pool allocation, and every
// block copies its part.  Passes of 32 channels leave room in LDS for three
// blocks per CU; a grid of fewer than 256 blocks (the coarse U-Net levels)
// takes 64-channel passes instead, half as many launches of a grid that
// cannot fill the chip anyway (measured faster on each layer of the 4-level
// U-Net).  Returns false when an allocation fails (the caller then runs the
// untiled kernel).
template <int KS, bool TR>
static bool conv_tile(const float *in, int batch, int cin, int h, int w, const float *wt, const float *bias, int cout,
                      float *out, int relu, hipStream_t st)
{
    const int PR = CONV_TROWS + KS - 1, PC = 64 + KS - 1;
    const long long P = (long long)h * w;
    const int nct = (cout + 31) / 32;
    const long long nblk = (long long)((w + 63) / 64) * ((h + CONV_TROWS - 1) / CONV_TROWS) * batch * nct;
    const int chunk = nblk < 256 ? 64 : 32;
    auto shape = [&](int nci, int &cinp, int &Kp, int &ps, int &wps, size_t &lds) {
        cinp = (nci + 7) / 8 * 8;
        Kp = (KS * KS * cinp + 31) / 32 * 32;
        conv_strides(cinp, Kp, ps, wps);
        lds = (size_t)PR * PC * ps * 2 + (size_t)32 * wps * 2 + (size_t)(Kp / 8) * 4;
    };
    int cinp, Kp, ps, wps;
    size_t lds;
    shape(std::min(cin, chunk), cinp, Kp, ps, wps, lds);   // the largest pass
    static bool raised = false;   // the 64 KiB default limit raised once per kernel
    if (lds > 65536 && !raised) {
        if (hipFuncSetAttribute(reinterpret_cast<const void *>(&k_conv_tile<KS>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) != hipSuccess)
            return false;
        raised = true;
    }
    void *wpk = nullptr;
    if (hipMallocAsync(&wpk, (size_t)nct * 32 * wps * 2, st) != hipSuccess) return false;
    dim3 grid((unsigned)((w + 63) / 64), (unsigned)((h + CONV_TROWS - 1) / CONV_TROWS), (unsigned)(batch * nct));
    for (int ci0 = 0; ci0 < cin; ci0 += chunk) {
        const int nci = std::min(chunk, cin - ci0);
        const bool last = ci0 + nci == cin;
        shape(nci, cinp, Kp, ps, wps, lds);
        const long long n = (long long)nct * 32 * wps;
        hipLaunchKernelGGL((k_conv_pack<KS, TR>), dim3(grid1(n)), dim3(256), 0, st, wt, ci0, nci, cin, cout, cinp,
                           wps, (__bf16 *)wpk, n);
        hipLaunchKernelGGL((k_conv_tile<KS>), grid, dim3(256), lds, st, in + ci0 * P, nci, h, w,
                           (const __bf16 *)wpk, last ? bias : nullptr, cout, out, last ? relu : 0, cinp, Kp,
                           (long long)cin * P, ci0 > 0 ? 1 : 0);
    }
    (void)hipFreeAsync(wpk, st);
    return true;
}

extern "C" int stfwi_conv2d_bf16(const float *in, int batch, int cin, int h, int w, const float *weight,
                                 const float *bias, int cout, int ksize, int relu, float *out, void *stream)
{
    if (!in || !weight || !out || batch < 1 || cin < 1 || cout < 1 || h < 1 || w < 1)
        return fail(STFWI_ERR_INVALID, "bad argument");
    if (ksize != 1 && ksize != 3) return fail(STFWI_ERR_UNSUPPORTED, "ksize=%d (use 1 or 3)", ksize);
    const long long P = (long long)h * w;
    if (P > (1LL << 30)) return fail(STFWI_ERR_INVALID, "image too large");
    hipStream_t st = (hipStream_t)stream;
    const bool tiled = ksize == 3 ? conv_tile<3, false>(in, batch, cin, h, w, weight, bias, cout, out, relu, st)
                                  : conv_tile<1, false>(in, batch, cin, h, w, weight, bias, cout, out, relu, st);
    dim3 grid((unsigned)((P + 255) / 256), (unsigned)((cout + 15) / 16), (unsigned)batch);
    if (tiled) {
    } else if (ksize == 3) {
        hipLaunchKernelGGL((k_conv_mfma<3, false>), grid, dim3(256), 0, st, in, cin, h, w, weight, bias, cout, out,
                           relu);
    } else {
        hipLaunchKernelGGL((k_conv_mfma<1, false>), grid, dim3(256), 0, st, in, cin, h, w, weight, bias, cout, out,
                           relu);
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int stfwi_conv2d_bf16_dgrad(const float *dout, int batch, int cout, int h, int w, const float *weight,
                                       int cin, int ksize, float *din, void *stream)
{
    if (!dout || !weight || !din || batch < 1 || cin < 1 || cout < 1 || h < 1 || w < 1)
        return fail(STFWI_ERR_INVALID, "bad argument");
    if (ksize != 1 && ksize != 3) return fail(STFWI_ERR_UNSUPPORTED, "ksize=%d (use 1 or 3)", ksize);
    const long long P = (long long)h * w;
    // a convolution of dout (cout channels) into cin channels with the
    // transposed, flipped weights
    hipStream_t st = (hipStream_t)stream;
    const bool tiled = ksize == 3 ? conv_tile<3, true>(dout, batch, cout, h, w, weight, nullptr, cin, din, 0, st)
                                  : conv_tile<1, true>(dout, batch, cout, h, w, weight, nullptr, cin, din, 0, st);
    dim3 grid((unsigned)((P + 255) / 256), (unsigned)((cin + 15) / 16), (unsigned)batch);
    if (tiled) {
    } else if (ksize == 3) {
        hipLaunchKernelGGL((k_conv_mfma<3, true>), grid, dim3(256), 0, st, dout, cout, h, w, weight, nullptr, cin,
                           din, 0);
    } else {
        hipLaunchKernelGGL((k_conv_mfma<1, true>), grid, dim3(256), 0, st, dout, cout, h, w, weight, nullptr, cin,
                           din, 0);
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" long long stfwi_conv2d_bf16_wgrad_work(int cin, int cout, int ksize)
{
    return (long long)WG_NPART * cout * cin * ksize * ksize;
}

extern "C" int stfwi_conv2d_bf16_wgrad(const float *in, const float *dout, int batch, int cin, int cout, int h, int w,
                                       int ksize, float *dw, float *db, int accumulate, float *work,
                                       long long work_elems, void *stream)
{
    if (!in || !dout || !dw || !work || batch < 1 || cin < 1 || cout < 1 || h < 1 || w < 1)
        return fail(STFWI_ERR_INVALID, "bad argument");
    if (ksize != 1 && ksize != 3) return fail(STFWI_ERR_UNSUPPORTED, "ksize=%d (use 1 or 3)", ksize);
    const long long K = (long long)cin * ksize * ksize, PT = (long long)batch * h * w;
    if (work_elems < stfwi_conv2d_bf16_wgrad_work(cin, cout, ksize))
        return fail(STFWI_ERR_INVALID, "work buffer holds %lld elements, needs %lld", work_elems,
                    stfwi_conv2d_bf16_wgrad_work(cin, cout, ksize));
    // up to WG_NPART partials, each over a strided set of 4 x 64 pixel tiles
    hipStream_t st = (hipStream_t)stream;
    const long long ntiles = (long long)batch * ((h + WG_TR - 1) / WG_TR) * ((w + 63) / 64);
    const int nz = (int)std::min<long long>(ntiles, WG_NPART);
    dim3 grid((unsigned)nz, (unsigned)((cin + 63) / 64), (unsigned)((cout + 31) / 32));
    if (ksize == 3)
        hipLaunchKernelGGL(k_conv_wgrad_tile<3>, grid, dim3(256), 0, st, in, cin, h, w, dout, cout, batch, work);
    else
        hipLaunchKernelGGL(k_conv_wgrad_tile<1>, grid, dim3(256), 0, st, in, cin, h, w, dout, cout, batch, work);
    const long long n = (long long)cout * K;
    hipLaunchKernelGGL(k_sum_chunks, dim3((unsigned)((n + 63) / 64)), dim3(256), 0, st, work, nz, n, dw, accumulate);
    if (db) {
        // bias gradient: per-chunk channel sums into work (after the
        // weight partials are summed), then summed in chunk order into db
        const long long bch = 4096;
        const int nb = (int)std::min<long long>((PT + bch - 1) / bch, 256);
        const long long pchb = (PT + nb - 1) / nb;
        hipLaunchKernelGGL(k_bias_grad, dim3(cout, nb), dim3(256), 0, st, dout, cout, batch, h * w, (int)pchb, work);
        hipLaunchKernelGGL(k_sum_chunks, dim3((unsigned)((cout + 63) / 64)), dim3(256), 0, st, work, nb, (long long)cout, db,
                           accumulate);
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int stfwi_relu_bwd(float *g, const float *y, long long n, void *stream)
{
    if (!g || !y || n < 1) return fail(STFWI_ERR_INVALID, "bad argument");
    hipLaunchKernelGGL(k_relu_bwd, dim3(grid1(n)), dim3(256), 0, (hipStream_t)stream, g, y, n);
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int stfwi_pool2(const float *in, long long planes, int h, int w, float *out, int backward, void *stream)
{
    if (!in || !out || planes < 1 || h < 2 || w < 2 || (h & 1) || (w & 1))
        return fail(STFWI_ERR_INVALID, "bad argument (planes of even size needed)");
    const long long n = backward ? planes * h * w : planes * (h / 2) * (w / 2);
    if (backward)
        hipLaunchKernelGGL(k_pool2_bwd, dim3(grid1(n)), dim3(256), 0, (hipStream_t)stream, in, planes, h, w, out);
    else
        hipLaunchKernelGGL(k_pool2, dim3(grid1(n)), dim3(256), 0, (hipStream_t)stream, in, planes, h, w, out);
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int stfwi_up2(const float *in, int batch, int planes, int h, int w, float *out, long long big_plane_stride,
                         long long small_batch_stride, long long big_batch_stride, int backward, void *stream)
{
    if (!in || !out || batch < 1 || planes < 1 || h < 1 || w < 1) return fail(STFWI_ERR_INVALID, "bad argument");
    const long long n = (long long)batch * planes * h * w * (backward ? 1 : 4);
    if (backward)   // in: gradient of the 2h x 2w planes, out: gradient of the h x w planes
        hipLaunchKernelGGL(k_up2_bwd, dim3(grid1(n)), dim3(256), 0, (hipStream_t)stream, in, planes, h, w, out,
                           big_plane_stride, batch, small_batch_stride, big_batch_stride);
    else            // in: h x w planes, out: 2h x 2w planes
        hipLaunchKernelGGL(k_up2, dim3(grid1(n)), dim3(256), 0, (hipStream_t)stream, in, planes, h, w, out,
                           big_plane_stride, batch, small_batch_stride, big_batch_stride);
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int stfwi_adam(float *p, const float *g, float *m, float *v, long long n, double lr, double b1, double b2,
                          double eps, int step, void *stream)
{
    if (!p || !g || !m || !v || n < 1 || step < 1) return fail(STFWI_ERR_INVALID, "bad argument");
    const float c1 = (float)(1.0 - std::pow(b1, step)), c2 = (float)(1.0 - std::pow(b2, step));
    hipLaunchKernelGGL(k_adam, dim3(grid1(n)), dim3(256), 0, (hipStream_t)stream, p, g, m, v, n, (float)lr, (float)b1,
                       (float)b2, (float)eps, c1, c2);
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int stfwi_mse(const float *y, const float *t, const float *wgt, long long n, double inv_wsum, float *g,
                         float *part, float *loss, void *stream)
{
    if (!y || !t || !g || !part || !loss || n < 1) return fail(STFWI_ERR_INVALID, "bad argument");
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_mse, dim3(256), dim3(256), 0, st, y, t, wgt, n, (float)inv_wsum, g, part);
    hipLaunchKernelGGL(k_sum_chunks, dim3(1), dim3(256), 0, st, part, 256, 1LL, loss, 0);
    HIP_TRY(hipGetLastError());
    return 0;
}
