/*
 * stfwi — MI355X (gfx950) elastic velocity-strain FWI gradient path.
 *
 * C-ABI of libstfwi.so.  All array arguments are DEVICE pointers owned by the
 * caller (the Python layer passes torch tensor data pointers), except where a
 * comment says "host".  Every call is asynchronous on the given hipStream_t
 * (passed as void*; NULL = the default stream) unless documented otherwise.
 * Functions return 0 on success and a negative stfwi_status on failure;
 * stfwi_last_error() then holds a message (thread-local).  No C++ exception
 * crosses this boundary.  A plan is not reentrant: use one plan per host
 * thread and stream.
 *
 * The reference (x-repos/stFWI) has no native boundary; its interface is the
 * Python call signatures cited per entry point below.
 */
#ifndef STFWI_H
#define STFWI_H

#ifdef __cplusplus
extern "C" {
#endif

#define STFWI_ABI_VERSION 2

typedef enum {
    STFWI_OK = 0,
    STFWI_ERR_INVALID = -1,     /* bad argument */
    STFWI_ERR_UNSUPPORTED = -2, /* fd_order / source type / mode not supported */
    STFWI_ERR_HIP = -3,         /* a HIP runtime call failed */
    STFWI_ERR_OOM = -4,         /* device allocation failed */
    STFWI_ERR_STATE = -5        /* call out of order (e.g. adjoint before forward) */
} stfwi_status;

enum { STFWI_F32 = 0, STFWI_F64 = 1 };
enum { STFWI_SRC_FORCE = 0, STFWI_SRC_MOMENT = 1 };
enum { STFWI_IMAGING_NONE = 0, STFWI_IMAGING_ZHOU = 1, STFWI_IMAGING_EXACT = 2 };
/* gradient parameter sets (the reference's invert='vs' | 'vs_vp' | 'all') */
enum { STFWI_GRAD_VS = 1, STFWI_GRAD_VS_VP = 2, STFWI_GRAD_ALL = 3 };
/* receiver channels, in the order of the reference's rec tuple
 * (vx, vz, exx, ezz), fwi/forward.py:465 */
enum { STFWI_CH_VX = 1, STFWI_CH_VZ = 2, STFWI_CH_EX = 4, STFWI_CH_EZ = 8 };
/* full-field snapshot selection, fwi/forward.py:450-460 ('*_full') */
enum { STFWI_SNAP_VX = 1, STFWI_SNAP_VZ = 2, STFWI_SNAP_EX = 4, STFWI_SNAP_EZ = 8, STFWI_SNAP_ES = 16 };

/* Grid, time axis and sizing.  Replaces the static arguments of
 * forward_jax (fwi/forward.py:119-123) and adjoint_jax
 * (fwi/adjoint_jax_kernel.py:61-69). */
typedef struct stfwi_config {
    int nz_dom, nx_dom;   /* physical grid (rows = z) */
    int pad;              /* C-PML thickness in cells */
    int free_surface;     /* 1: free surface on top, PML on 3 sides */
    int fd_order;         /* 2, 4 or 8 */
    double dx, dz, dt;
    double fc;            /* PML alpha tuning frequency */
    int dtype;            /* STFWI_F32 or STFWI_F64 */
    int n_shots_max;      /* shots per batched launch */
    int nt_max;           /* longest record */
    int nrec_max;         /* most receivers of any shot */
    int imaging;          /* STFWI_IMAGING_* (history storage is sized for it) */
    int grad_params;      /* STFWI_GRAD_* */
    int store_history;    /* 1: keep the strain history for the adjoint */
    int ckpt_every;       /* 0: history of every step; K > 0: checkpoint the
                           * forward state every K steps and recompute each
                           * K-step segment's history during the adjoint
                           * (memory: nt/K states + K+1 history frames; the
                           * reference's jax.checkpoint blocks,
                           * fwi/forward.py:471-501).  Results do not depend
                           * on K. */
    long long hist_frames;  /* history capacity in frames summed over the
                           * shots of one stfwi_set_shots (each shot keeps
                           * its own nt frames, so uneven records do not pay
                           * for the longest); 0 = n_shots_max * nt_max.
                           * Ignored with ckpt_every > 0 (K + 1 frames a
                           * shot). */
} stfwi_config;

/* One shot.  Indices are physical-domain (the reference's src_x, src_z,
 * rec_x, rec_z, fwi/forward.py:240-242). */
typedef struct stfwi_shot {
    int src_x, src_z;
    int src_type;         /* STFWI_SRC_* */
    int nt;               /* record length, <= nt_max */
    int rec_offset;       /* first receiver in the rec_x/rec_z host arrays */
    int nrec;             /* <= nrec_max */
} stfwi_shot;

typedef struct stfwi_plan stfwi_plan;

int stfwi_abi_version(void);
const char *stfwi_last_error(void);

/* Allocate device workspace (state, PML strips, history).  Host call,
 * synchronous.  Replaces the per-call set-up of forward_jax
 * (fwi/forward.py:167-247). */
int stfwi_plan_create(const stfwi_config *cfg, stfwi_plan **out);
int stfwi_plan_destroy(stfwi_plan *plan);
/* bytes of device memory the plan holds (host call) */
long long stfwi_plan_bytes(const stfwi_plan *plan);

/* Bytes of device memory stfwi_plan_create would allocate for cfg, without
 * allocating (host call; no reference counterpart). */
int stfwi_plan_footprint(const stfwi_config *cfg, long long *bytes);

/* Checkpoint planner (no reference counterpart; the reference fixes its
 * jax.checkpoint block size by hand, fwi/forward.py:471-501,
 * Main_strain_fwi.py:42-46): *ckpt_every receives 0 when the plan of cfg
 * with the whole history fits in budget_bytes, otherwise the segment
 * length K whose plan has the smallest footprint (about sqrt(nt_max) state
 * checkpoints plus one segment of history).  STFWI_ERR_OOM when no K fits;
 * STFWI_ERR_UNSUPPORTED when checkpointing would be needed but the imaging
 * mode does not allow it (Zhou imaging of rho).  cfg->ckpt_every is
 * ignored. */
int stfwi_plan_choose_ckpt(const stfwi_config *cfg, long long budget_bytes, int *ckpt_every);

/* Model: vs, vp, rho (nz_dom x nx_dom, plan dtype).  Builds the padded
 * moduli, buoyancy and the 1-D C-PML profiles on the device
 * (fwi/forward.py:193-238, fwi/forward.py:35-104).  b_img (optional, may be
 * NULL) is the physical buoyancy the Zhou imaging uses (fwi/imaging.py:113-114
 * computes it in the model's storage dtype); NULL means 1/rho. */
int stfwi_set_model(stfwi_plan *plan, const void *vs, const void *vp, const void *rho,
                    const void *b_img, void *stream);

/* Shots: host table of n_shots, host receiver index arrays, device wavelets
 * (n_shots x nt_max, plan dtype).  Host call: uploads the table and the
 * collision-merged receiver cells (duplicates accumulate, as the reference's
 * .at[].add does, fwi/adjoint_jax_kernel.py:203-206). */
int stfwi_set_shots(stfwi_plan *plan, const stfwi_shot *shots, int n_shots,
                    const int *rec_x, const int *rec_z, const void *wavelets, void *stream);

/* Forward modelling of all shots in one batched time loop
 * (forward_jax, fwi/forward.py:119-501).
 *   rec_out   n_shots x 4 x nt_max x nrec_max, channels vx vz exx ezz (or NULL)
 *   rec_mask  STFWI_CH_* bits to record
 *   illum     n_shots x nz_t x nx_t accumulators of sum_t field^2 (or NULL),
 *             illum_comp = 0..3 (vx, vz, exx, ezz) as in
 *             fwi/illumination.py:22-34; the caller zeroes it
 *   snap      n_shots x popcount(snap_mask) x nt_max x nz_t x nx_t full
 *             padded fields after each step (or NULL), STFWI_SNAP_* bits
 * With store_history the plan keeps the strains needed by stfwi_adjoint. */
int stfwi_forward(stfwi_plan *plan, void *rec_out, int rec_mask, void *illum, int illum_comp,
                  void *snap, int snap_mask, void *stream);

/* Hand-coded adjoint of all shots (adjoint_jax,
 * fwi/adjoint_jax_kernel.py:65-483) with the imaging condition of the plan
 * fused into each reverse step.
 *   rec_adj   n_shots x 4 x nt_max x nrec_max receiver adjoints
 *   src_adj   n_shots x nt_max source-wavelet adjoint (or NULL)
 *   snap      n_shots x popcount(snap_mask) x nt_max x nz_t x nx_t adjoint
 *             carry-out fields (or NULL; fwi/adjoint_jax_kernel.py:470-476)
 * Gradients stay in the plan until stfwi_gradient_finalize. */
int stfwi_adjoint(stfwi_plan *plan, const void *rec_adj, void *src_adj, void *snap, int snap_mask,
                  void *stream);

/* Sum the per-shot accumulators and chain them to (Vs, Vp, rho) on the
 * physical grid (fwi/imaging.py:164-167 for Zhou; the exact chain for
 * 'ad').  vs/vp/rho are the model arrays whose dtype (chain_dtype,
 * STFWI_F32/F64) the chain factors are formed in, as the reference forms
 * 2*rho*Vs in the model's own dtype.  g_out: 3 x nz_dom x nx_dom in the plan
 * dtype; parameters outside grad_params are written as zeros. */
int stfwi_gradient_finalize(stfwi_plan *plan, const void *vs, const void *vp, const void *rho,
                            int chain_dtype, void *g_out, void *stream);

/* Residual prep of all shots (fwi/gradient.py:139-156, fwi/filters.py):
 * taper the last n_taper samples, filter with the DF-I biquad cascade sos
 * (host, n_sos x 6, may be NULL), misfit (scale/n_shots) mean((F T pred -
 * F T obs)^2) per shot into loss_out (n_shots, plan dtype), and the receiver
 * cotangent into channel `component` of rec_adj (other channels zeroed).
 *   pred      n_shots x 4 x nt_max x nrec_max (the forward's rec_out)
 *   obs       n_shots x nt_max x nrec_max
 *   n_shots_norm  the n_shots of the reference's normalisation */
int stfwi_residual(stfwi_plan *plan, const void *pred, const void *obs, int component,
                   const double *sos, int n_sos, int n_taper, double scale, int n_shots_norm,
                   void *loss_out, void *rec_adj, void *stream);

/* Gaussian smoothing of a (nz, nx) array: normalised kernel of radius
 * int(3 sigma + 0.5), 'same'-mode zero padding (gaussian_smooth_2d,
 * fwi/loss.py:17-30).  tmp is an (nz, nx) workspace not aliasing in/out. */
int stfwi_smooth2d(const void *in, void *tmp, void *out, int nz, int nx, double sigma, int dtype, void *stream);

/* Illumination multiplier H = 1/(Ws + eps max Ws), normalised to max 1, from
 * n_shots padded energies ws (n_shots x nz_t x nx_t) summed over shots and
 * cropped to rows [z0, z0+nz_c), columns [pad, pad+nx_c)
 * (compute_illumination_for_stage, fwi/illumination.py:38-48). */
int stfwi_illumination(const void *ws, int n_shots, int nz_t, int nx_t, int z0, int pad, int nz_c, int nx_c,
                       double eps, int dtype, void *out, void *stream);

/* Plessix-Mulder multiplier from summed padded energies: We (eprecond3,
 * fwi/preconditioning.py:65-139) into we_pad (nz_t x nx_t), then the
 * normalised physical-grid multiplier 1/We / max into out (nz_dom x nx_dom),
 * as fwi/optim_lbfgsb.py:90-100.  rec_x/rec_z are host arrays. */
int stfwi_eprecond3(const void *ws, int n_shots, int nz_t, int nx_t, int free_surface, int pad, int nz_dom,
                    int nx_dom, const int *rec_x, const int *rec_z, int nrec, double dx, double dz, double eps,
                    int dtype, void *we_pad, void *out, void *stream);

/* Shin energy weight We = sqrt(Ws Wr) + eps max(sqrt(Ws Wr)) over n elements
 * (eprecond1, fwi/preconditioning.py:37-62); dtype of ws, wr and we. */
int stfwi_eprecond1(const void *ws, const void *wr, long long n, double eps, int dtype, void *we, void *stream);

/* Allocation-free variant of stfwi_eprecond1 (no counterpart in the
 * reference; for callers that cannot allocate device memory: code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results, bit for bit.  Its reduction scratch (the
 * partial maxima and the maximum, 1025 doubles = 8 KiB) is a file-scope
 * device array of the library, one copy per GPU, held until the process
 * ends; any n fits it.  Calls on one GPU share that copy: calls that
 * overlap on one GPU (other streams or threads) race. */
int stfwi_eprecond1_noalloc(const void *ws, const void *wr, long long n, double eps, int dtype, void *we,
                            void *stream);

/* Depth taper of the gradient, FP64 (nz x nx), rows as taper_grad
 * (fwi/preconditioning.py:144-183): 0 above gradt1, exp(-(gradt2-j)^2) up to
 * gradt2, (j dz)^exp_taper below; with bottom != 0 the roll-off between gradb1
 * and gradb2 and 0 below gradb2. */
int stfwi_taper_grad(int nz, int nx, double dz, int gradt1, int gradt2, double exp_taper, int bottom, int gradb1,
                     int gradb2, double *out, void *stream);

/* Source taper, FP64 (nz x nx), as build_source_taper (fwi/taper.py:16-84):
 * the erf patch of radius srt_radius stamped (minimum) around every source,
 * normalised to [0, 1], then a (2 filt_size + 1)^2 zero box at each source.
 * src_xz is a host array of nsrc (x, z) grid-index pairs.  Host-synchronous. */
int stfwi_source_taper(int nz, int nx, double dh, const int *src_xz, int nsrc, double srt_radius, int filt_size,
                       double *out, void *stream);

/* In-place trace filtering x[t][trace] (nt x ntr): taper the last n_taper
 * samples then the DF-I biquad cascade (jax_sosfilt / cosine_taper_end,
 * fwi/filters.py:14-42), or with transpose=1 the exact transpose of that
 * map.  sos is a host (n_sos x 6) array, n_sos <= 8. */
int stfwi_sosfilt(void *x, int nt, int ntr, const double *sos, int n_sos, int n_taper, int transpose, int dtype,
                  void *stream);

/* Zhou imaging of given history cubes (imaging_condition,
 * fwi/imaging.py:83-169): ex, ez, es, ua, va (and optionally vx, vz, both or
 * neither) are nt x nz x nx padded-grid cubes, b_full the padded buoyancy
 * (nz x nx); out receives g_lambda, g_mu, g_rho_lame on the padded grid
 * (3 x nz x nx). */
int stfwi_zhou_imaging(const void *ex, const void *ez, const void *es, const void *ua, const void *va,
                       const void *b_full, const void *vx, const void *vz, int nt, int nz, int nx, double dx,
                       double dz, double dt, int dtype, void *out, void *stream);

/* Diagnostics (no reference counterpart): time `reps` back-to-back launches
 * of one time-step kernel over all shots of the plan (which: 0 forward
 * velocity, 1 forward strain, 2 adjoint strain, 3 adjoint velocity with
 * imaging) with HIP events on `stream`; *ms receives the total.  With
 * grouped != 0 each repetition launches every shot group of the plan on its
 * own stream, as the time loops run them; otherwise one launch covers all
 * shots.  Host call, synchronous.  It overwrites the plan's wavefield state,
 * so a forward or adjoint that ran before it has to run again before the
 * next adjoint or finalize. */
int stfwi_time_kernel(stfwi_plan *plan, int which, int reps, int grouped, void *stream, double *ms);

/* Tile-shape autotuner (no reference counterpart): for each of the four
 * time-step kernels time `reps` launches at every compiled tile height
 * (4/8/16/32 rows of 64 columns in FP32, 4/8 in FP64) on the plan's shots,
 * each repetition launching every shot group (stfwi_set_groups) on its own
 * stream as the time loops do, in three interleaved rounds of which the
 * fastest counts, and keep the fastest height.  tz_out (4 ints) and us_out (4 x 4 average
 * microseconds, 0 where not compiled) may be NULL.  Host call, synchronous; overwrites the
 * wavefield state, so run it before stfwi_forward. */
int stfwi_autotune(stfwi_plan *plan, int reps, void *stream, int *tz_out, double *us_out);

/* Set the four tile heights explicitly (kernel order: forward velocity,
 * forward strain, adjoint strain, adjoint velocity). */
int stfwi_set_tiles(stfwi_plan *plan, const int *tz);

/* Graph replay of the time loops (no reference counterpart): refused.
 * enable != 0 returns STFWI_ERR_UNSUPPORTED, because a host segmentation
 * fault with replay on has no cause shown yet; enable = 0 keeps the eager
 * launches, the only path. */
int stfwi_set_graphs(stfwi_plan *plan, int enable);

/* Adjoint wavefield energy (no reference counterpart; the Wr that eprecond1,
 * fwi/preconditioning.py:37-62, takes): while energy is non-NULL every
 * stfwi_adjoint adds, at every reverse step, the square of adjoint field comp
 * (0 U_a, 1 V_a, 2 exx_a, 3 ezz_a, 4 exz_a; the values snapshots record) to
 * energy[shot][nz_t][nx_t] on the padded grid.  The caller zeroes it. */
int stfwi_set_adjoint_energy(stfwi_plan *plan, void *energy, int comp);

/* Split the shots into `groups` (1..4) contiguous groups whose time loops
 * run concurrently on the plan's own streams (default priority), joined
 * back into the caller's stream; results do not depend on the grouping. */
int stfwi_set_groups(stfwi_plan *plan, int groups);

/* No reference counterpart (a diagnostic).  Each time-step kernel has a
 * specialised instantiation for runs without its optional outputs (velocity
 * history, illumination energy, snapshots, adjoint energy, source-wavelet
 * adjoint) and, for the velocity adjoint, with exact imaging of Vs only;
 * the launches pick it when the run's arguments match.  enable = 0 makes
 * every launch use the general instantiation, whose results are the same
 * bit for bit.  Default 1. */
int stfwi_set_lean(stfwi_plan *plan, int enable);

/* ---- L-BFGS in float64 on the device (Nocedal & Wright Alg. 7.4).  The
 * reference runs its quasi-Newton step on the host: scipy L-BFGS-B in
 * run_multiscale_fwi (fwi/optim_lbfgsb.py:409-434) and jaxopt.LBFGS in
 * run_lbfgs (fwi/optimizer.py:110-117).  All vectors are device float64;
 * each call is one workgroup with a fixed summation order (bitwise
 * reproducible). ---- */

/* d = -H g from m stored pairs: rows slots[0..m) (oldest first) of S and Y
 * (row stride ld), rho[k] = 1/(y.s) of row slots[k] (host arrays, m <= 64),
 * H0 = gamma I.  *gd (device scalar, may be NULL) receives g . d. */
int stfwi_lbfgs_direction(const double *g, const double *S, const double *Y, const double *rho, const int *slots,
                          int m, long long n, long long ld, double gamma, double *d, double *gd, void *stream);

/* s = x_new - x_old, y = g_new - g_old into s_out / y_out; dots (device, 3)
 * receives y.s, y.y, s.s. */
int stfwi_lbfgs_pair(const double *x_new, const double *x_old, const double *g_new, const double *g_old,
                     long long n, double *s_out, double *y_out, double *dots, void *stream);

/* *out = a . b (device scalar). */
int stfwi_dot64(const double *a, const double *b, long long n, double *out, void *stream);

/* out = clip(x + t d, lo, hi) elementwise (the projection onto per-cell
 * bounds of fwi/optim_lbfgsb.py:298-309); lo / hi may be NULL. */
int stfwi_step_clip(const double *x, const double *d, double t, const double *lo, const double *hi, long long n,
                    double *out, void *stream);

/* v[i] = 0 where x[i] sits on a bound and g[i] points out of the box
 * (x <= lo, g > 0 or x >= hi, g < 0): the binding set masked out of the
 * gradient and the search direction of the projected method. */
int stfwi_bound_mask(const double *x, const double *g, const double *lo, const double *hi, long long n, double *v,
                     void *stream);

/* ---- Learned gradient preconditioner (no reference counterpart; the
 * reference's plain method has no learned part).  A small U-Net whose
 * convolutions run on the matrix cores: operands rounded to BF16, FP32
 * accumulation.  Activations are FP32 NCHW (batch x channels x h x w),
 * weights FP32 [cout][cin][k][k], k = 1 or 3, stride 1, zero 'same'
 * padding.  All pointers are device pointers. ---- */

/* out = conv(in, weight) + bias, then max(., 0) if relu. */
int stfwi_conv2d_bf16(const float *in, int batch, int cin, int h, int w, const float *weight, const float *bias,
                      int cout, int ksize, int relu, float *out, void *stream);
/* din = gradient of the convolution's input from dout (transposed,
 * flipped weights). */
int stfwi_conv2d_bf16_dgrad(const float *dout, int batch, int cout, int h, int w, const float *weight, int cin,
                            int ksize, float *din, void *stream);
/* dw (and db if not NULL) = gradient of the weights (bias) from in and dout,
 * summed over the batch; accumulate = 1 adds to dw / db.  work is a scratch
 * buffer of stfwi_conv2d_bf16_wgrad_work(cin, cout, ksize) floats. */
long long stfwi_conv2d_bf16_wgrad_work(int cin, int cout, int ksize);
int stfwi_conv2d_bf16_wgrad(const float *in, const float *dout, int batch, int cin, int cout, int h, int w, int ksize,
                            float *dw, float *db, int accumulate, float *work, long long work_elems, void *stream);
/* g *= (y > 0) */
int stfwi_relu_bwd(float *g, const float *y, long long n, void *stream);
/* 2 x 2 average pooling of `planes` h x w planes (h, w even) or, with
 * backward = 1, its gradient (in: h/2 x w/2 gradients, out: h x w). */
int stfwi_pool2(const float *in, long long planes, int h, int w, float *out, int backward, void *stream);
/* Nearest x2 upsampling of h x w planes into 2h x 2w planes laid out with
 * big_plane_stride (so it can fill part of a concatenation buffer), or with
 * backward = 1 its gradient (in: the 2h x 2w gradients, out: h x w). */
int stfwi_up2(const float *in, int batch, int planes, int h, int w, float *out, long long big_plane_stride,
              long long small_batch_stride, long long big_batch_stride, int backward, void *stream);
/* Adam step (Kingma & Ba, Alg. 1) of step number `step` >= 1 on FP32 p. */
int stfwi_adam(float *p, const float *g, float *m, float *v, long long n, double lr, double b1, double b2, double eps,
               int step, void *stream);
/* Weighted MSE: g = 2 w (y - t) inv_wsum, *loss = sum w (y - t)^2 (device
 * scalar; multiply by inv_wsum for the mean); part is 256 floats of
 * scratch; wgt may be NULL (weight 1). */
int stfwi_mse(const float *y, const float *t, const float *wgt, long long n, double inv_wsum, float *g, float *part,
              float *loss, void *stream);

#ifdef __cplusplus
}
#endif
#endif
