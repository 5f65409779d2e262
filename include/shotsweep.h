/*
 * shotsweep.h - C ABI of libshotsweep.so, the gfx950 HIP implementation of the
 * batched 2-D constant-density acoustic finite-difference simulator, its exact
 * discrete adjoint and the per-shot RTM imaging condition.
 *
 * Reference interface replaced: the PyTorch op sequences inside
 * openfwi/physics/gpu.py (class Sim).  Every entry point names the lines it
 * replaces.  The Python module shotsweep_amd.gpu keeps the reference's call
 * surface (Sim, lap, lap_adjoint, ...) and calls these functions through
 * ctypes on raw device pointers.
 *
 * Conventions
 *  - All arrays are device pointers, C-contiguous, in the element type named
 *    by the suffix (_f32 = float, _f64 = double).  The caller allocates every
 *    buffer; the library never frees caller memory and keeps no pointer after
 *    a call returns.
 *  - Padded grid: H = nz + 2*pad rows, W = nx + 2*pad columns.  A
 *    "propagation" is one (row b, shot s) pair; there are nb*ns of them.
 *  - Calls are asynchronous and ordered on `stream` (a hipStream_t; NULL is the
 *    default stream).  They return 0 on success and a negative code on error;
 *    ss_last_error() then gives a message (thread-local).  No C++ exception
 *    crosses the ABI.
 *  - Arithmetic reproduces the reference's FP32/FP64 operation order (no
 *    contraction of separately rounded operations into FMAs, correctly
 *    rounded division), so results equal the reference CPU path bit for bit
 *    except where noted (the reduction order of sums).
 */
#ifndef SHOTSWEEP_H
#define SHOTSWEEP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SS_MAX_SHOTS 32

/* Error codes */
#define SS_OK 0
#define SS_EINVAL (-1)   /* bad argument (shape, range, null pointer) */
#define SS_EHIP (-2)     /* HIP runtime error */
#define SS_EUNSUP (-3)   /* requested implementation cannot run this shape */

/* Implementation selector for the time-loop entry points. */
#define SS_IMPL_AUTO 0      /* resident when the plane fits the LDS, else stepwise (ss_rtm_*: tiled
                               checkpoint segments for batches of at most CUs/4 propagations or
                               planes that do not fit the LDS) */
#define SS_IMPL_STEPWISE 1  /* one launch per time step, planes in HBM (FP32/FP64, any size) */
#define SS_IMPL_RESIDENT 2  /* one persistent launch per step range, plane in LDS (FP32) */
#define SS_IMPL_TILED 3     /* ss_rtm_* checkpoint segments on the tiled sweep kernels (FP32) */
#define SS_IMPL_UNTILED 4   /* ss_rtm_*: as SS_IMPL_AUTO, but never the tiled kernels */

/*
 * Geometry of one batch of propagations (openfwi/physics/constants.py:4-7,
 * gpu.py:14, generalised to rectangular grids and any shot count).
 */
typedef struct ss_geom {
  int32_t nb;                 /* velocity rows (models x backgrounds), gpu.py:126 len(v) */
  int32_t ns;                 /* shots per row, gpu.py:95 (5) */
  int32_t nz, nx;             /* model grid, constants.py:6 (70 x 70) */
  int32_t pad;                /* absorbing-layer width in cells, gpu.py:77 (40) */
  int32_t nt;                 /* time steps, constants.py:6 (1000) */
  int32_t row;                /* model row of sources and receivers, constants.py:7 (row 1) */
  int32_t src[SS_MAX_SHOTS];  /* source columns in model coordinates, gpu.py:14 */
} ss_geom;

/* Message of the last failed call on this thread ("" if none). */
const char *ss_last_error(void);

/* ABI version (incremented on any signature change). */
int ss_version(void);

/* Largest padded plane (H*W cells) the resident FP32 kernels hold on one CU. */
int64_t ss_resident_max_cells(void);

/*
 * Bytes of the workspace that ss_rtm_* needs for one gradient or images call
 * on g (the forward with its tape, residual, adjoint and fold), with elements
 * of elem_size bytes (4 or 8): segment >= nt keeps the whole Laplacian
 * history; a smaller segment keeps checkpoints every `segment` steps plus two
 * alternating segment buffers.  The same size serves every implementation
 * selector.  This implementation's counterpart of gpu.py:70-73
 * (imaging_memory_bytes), which estimates the reference's own scheme.
 * Returns -1 (ss_last_error set) on an invalid geometry or size.  Host-only
 * arithmetic: no device is touched.
 */
int64_t ss_workspace_bytes(const ss_geom *g, int segment, int elem_size);

/*
 * Forward time loop over steps [t0, t1), replacing gpu.py:111-115 (Sim._step)
 * and the loop gpu.py:126-135 (Sim.forward), including lap (gpu.py:33-37).
 *
 *   c2      (nb, H, W)   vp^2 (dt/dx)^2, formed in float64 and rounded (gpu.py:107)
 *   A, B    (H, W)       sponge coefficients 1/(1+a), 1-a (gpu.py:85-90)
 *   w       (nt)         wavelet * 1e3 (gpu.py:94)
 *   amp     (nb, ns)     (v[row, src]/2000)^2 (gpu.py:108)
 *   p0, p1  (nb, ns, H, W) state before step t0 (in) and after step t1-1 (out):
 *                          p0 <- field at t1-1, p1 <- field at t1
 *   gathers (nb, ns, nt, nx) or NULL: receiver samples of steps [t0, t1) (gpu.py:134)
 *   lap_hist (nb, ns, t1-t0, H, W) or NULL: Laplacian of each step (gpu.py:112, 172-175)
 *   ckpt    (nb, ns, ceil((t1-t0)/ckpt_every), 2, H, W) or NULL: state (p0, p1)
 *           before every step t with (t - t0) % ckpt_every == 0 (gpu.py:131-132)
 */
int ss_forward_f32(const ss_geom *g, int t0, int t1, const float *c2, const float *A, const float *B,
                   const float *w, const float *amp, float *p0, float *p1, float *gathers,
                   float *lap_hist, float *ckpt, int ckpt_every, int impl, void *stream);
int ss_forward_f64(const ss_geom *g, int t0, int t1, const double *c2, const double *A, const double *B,
                   const double *w, const double *amp, double *p0, double *p1, double *gathers,
                   double *lap_hist, double *ckpt, int ckpt_every, int impl, void *stream);

/*
 * Adjoint time loop over steps t1-1 down to t0, replacing the reverse loop
 * gpu.py:177-184 (adjoint recursion with lap_adjoint gpu.py:40-53, receiver
 * injection, imaging condition, source-amplitude term, illumination).
 *
 *   sd       (nb, ns)   source-amplitude derivative 2 v[row,src]/2000^2 (gpu.py:168)
 *   res      (nb, ns, nt, nx) muted residual (gpu.py:158-160)
 *   lap_hist (nb, ns, t1-t0, H, W) Laplacians of the forward field at steps [t0, t1)
 *   m1, m2   (nb, ns, H, W) adjoint state (in at t1, out at t0): m1 <- mu(t0), m2 <- mu(t0+1)
 *   gacc     (nb, ns, H, W) running sum of A*mu*L (gpu.py:181)
 *   illum    (nb, ns, nz, nx) running sum of L^2 over the model cells (gpu.py:183)
 *   src_grad (nb, ns)  running source-amplitude term (gpu.py:182)
 */
int ss_adjoint_f32(const ss_geom *g, int t0, int t1, const float *c2, const float *A, const float *B,
                   const float *w, const float *sd, const float *res, const float *lap_hist,
                   float *m1, float *m2, float *gacc, float *illum, float *src_grad, int impl, void *stream);
int ss_adjoint_f64(const ss_geom *g, int t0, int t1, const double *c2, const double *A, const double *B,
                   const double *w, const double *sd, const double *res, const double *lap_hist,
                   double *m1, double *m2, double *gacc, double *illum, double *src_grad, int impl,
                   void *stream);

/*
 * One gradient or images call on the device, replacing the body of
 * Sim._gradient and Sim.images (gpu.py:152-188, 196-210): the forward with its
 * tape (gpu.py:117-136, 154-157), the muted residual (gpu.py:158-160), the
 * exact adjoint over the whole Laplacian history or over checkpoint segments
 * recomputed from their checkpoints (gpu.py:161-184), the fold with the
 * source-amplitude term (gpu.py:185-187) and, when img is given, the image
 * epilogue (gpu.py:206-209).  Every intermediate lives in `workspace`
 * (256-byte aligned, at least ss_workspace_bytes(g, segment, elem size)
 * bytes); the library allocates no device memory for it.  The two stages may
 * run in separate calls on one workspace: SS_RTM_FORWARD leaves the tape and
 * the gathers there, a later SS_RTM_ADJOINT with the same geometry, segment
 * and impl reads them (an autograd backward passes its data-space gradient
 * as `seed`).  Checkpoint segments recompute segment k-1 on a side stream
 * owned by the library (one per thread and device) while segment k's adjoint
 * runs on `stream`; results do not depend on that overlap.  Calls that share
 * a workspace must not overlap in time.
 */
#define SS_RTM_FORWARD 1
#define SS_RTM_ADJOINT 2

typedef struct ss_rtm_args {
  int32_t stages;      /* SS_RTM_FORWARD, SS_RTM_ADJOINT or both */
  int32_t segment;     /* >= nt: whole Laplacian history; else checkpoints every `segment` steps (gpu.py:131) */
  int32_t impl;        /* SS_IMPL_* for the time loops */
  int32_t obs_repeat;  /* row b of the batch uses row b / obs_repeat of obs */
  int32_t laplacian;   /* image: 1 applies the Laplacian filter (gpu.py:208-209) */
  int32_t timing;      /* 1: HIP events around every time-loop launch (read with ss_rtm_timing); the
                          segment recomputes then run on `stream` so each interval is one launch */
  double dt, dx;       /* time step (s) and grid spacing (m) of the mute (gpu.py:146-149) */
  double t_mute, ramp; /* mute delay and ramp length (s): 0.12 and 0.03 in the reference */
  double eps;          /* image: illumination regularisation (gpu.py:206) */
  double k;            /* (dt/dx)^2, rounded to the working type (gpu.py:185) */
  const void *c2, *A, *B, *w, *amp; /* as in ss_forward_* */
  const void *sd;      /* (nb, ns) source-amplitude derivative, as in ss_adjoint_* */
  const void *v;       /* (nb, nz, nx) velocity of the rows, as in ss_gradient_* */
  const void *obs;     /* (nb / obs_repeat, ns, nt, nx) observed gathers, or NULL when seed is given */
  const void *v_mute;  /* (nb) mute velocities, or NULL: no mute */
  const void *seed;    /* (nb, ns, nt, nx) data-space seed used instead of the residual, or NULL */
  void *gathers;       /* out (nb, ns, nt, nx) simulated gathers, or NULL */
  void *grad, *illum;  /* out (nb, ns, nz, nx): per-shot gradient and illumination (adjoint stage) */
  void *img;           /* out (nb, ns, nz, nx) image, or NULL (adjoint stage) */
} ss_rtm_args;

int ss_rtm_f32(const ss_geom *g, const ss_rtm_args *a, void *workspace, int64_t workspace_bytes, void *stream);
int ss_rtm_f64(const ss_geom *g, const ss_rtm_args *a, void *workspace, int64_t workspace_bytes, void *stream);

/*
 * ss_rtm_f32 without a workspace argument and without any device allocation:
 * no counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The workspace is a fixed 256 MiB device array inside the
 * library (one copy per GPU, held until the process ends); a call whose
 * ss_workspace_bytes(g, a->segment, 4) exceeds it fails with SS_EINVAL.  Every
 * launch is a direct launch on `stream` (no side stream, no graphs, no
 * autotuner trials, no timing events: a->timing is ignored).  All calls on one
 * GPU share the array and it keeps the last call's contents: calls that
 * overlap on one GPU race, and a forward-only call's tape survives only until
 * the next call on that GPU.
 */
int ss_rtm_noalloc_f32(const ss_geom *g, const ss_rtm_args *a, void *stream);

/* Launch counts and summed durations (ms) of the forward and adjoint time-loop
 * launches of this thread's last ss_rtm_* call with timing = 1 (waits for them). */
int ss_rtm_timing(int *nforward, double *forward_ms, int *nadjoint, double *adjoint_ms);

/*
 * Residual with the direct-wave mute, replacing gpu.py:142-150 and 158-160:
 *   res = (sim - obs) * mute  (mute omitted when v_mute is NULL)
 *   sim      (nb, ns, nt, nx)
 *   obs      (nb / obs_repeat, ns, nt, nx): row b of sim uses row b / obs_repeat of obs
 *            (obs_repeat = 1 reproduces the reference's explicit np.repeat)
 *   v_mute   (nb) mute velocity per row, or NULL
 *   dt_s, dx_m, t_mute, ramp: dt (s), dx (m), mute delay (s), ramp length (s, 0.03)
 *   res      (nb, ns, nt, nx) output (may alias sim)
 * ss_mute_weights_* writes the weights alone, (nb, ns, nt, nx).
 */
int ss_residual_f32(const ss_geom *g, const float *sim, const float *obs, int obs_repeat,
                    const float *v_mute, double dt_s, double dx_m, double t_mute, double ramp,
                    float *res, void *stream);
int ss_residual_f64(const ss_geom *g, const double *sim, const double *obs, int obs_repeat,
                    const double *v_mute, double dt_s, double dx_m, double t_mute, double ramp,
                    double *res, void *stream);
int ss_mute_weights_f32(const ss_geom *g, const float *v_mute, double dt_s, double dx_m, double t_mute,
                        double ramp, float *out, void *stream);
int ss_mute_weights_f64(const ss_geom *g, const double *v_mute, double dt_s, double dx_m, double t_mute,
                        double ramp, double *out, void *stream);

/*
 * Gradient epilogue, replacing gpu.py:185-187 and fold_edge_padding
 * gpu.py:56-67:  grad = fold(gacc * 2 * vp * k) ; grad[row, src] += src_grad
 *   v        (nb, nz, nx) model velocity in the working type (vp = edge padding of v)
 *   k        (dt/dx)^2 rounded to the working type
 *   grad     (nb, ns, nz, nx) output
 * Sums over the padding strips run in ascending index order; the reference's
 * torch reduction order may differ in the last bit.
 */
int ss_gradient_f32(const ss_geom *g, const float *gacc, const float *v, const float *src_grad, float k,
                    float *grad, void *stream);
int ss_gradient_f64(const ss_geom *g, const double *gacc, const double *v, const double *src_grad, double k,
                    double *grad, void *stream);

/*
 * Image epilogue, replacing gpu.py:206-209:
 *   den = illum + eps * max(illum) per propagation ; img = -grad / max(den, tiny)
 *   if laplacian: img = -lap(replicate_pad2(img)) on the model grid
 *   grad, illum (nb, ns, nz, nx) ; img (nb, ns, nz, nx) output
 */
int ss_image_f32(const ss_geom *g, const float *grad, const float *illum, float eps, int laplacian,
                 float *img, void *stream);
int ss_image_f64(const ss_geom *g, const double *grad, const double *illum, double eps, int laplacian,
                 double *img, void *stream);

/*
 * Tangent-linear (Born) forward over steps [t0, t1): no reference counterpart
 * (the reference has no linearised operator); it is the exact derivative of
 * ss_forward_* with respect to velocity, used for the dot-product test.
 *   dc2   (nb, H, W)  2 vp dvp (dt/dx)^2 ;  damp (nb, ns) 2 v dv / 2000^2 at the source
 *   q0,q1 (nb, ns, H, W) perturbation state ; p0,p1 background state (as ss_forward_*)
 *   born  (nb, ns, nt, nx) receiver samples of the perturbation
 */
int ss_born_f32(const ss_geom *g, int t0, int t1, const float *c2, const float *dc2, const float *A,
                const float *B, const float *w, const float *amp, const float *damp, float *p0, float *p1,
                float *q0, float *q1, float *born, void *stream);
int ss_born_f64(const ss_geom *g, int t0, int t1, const double *c2, const double *dc2, const double *A,
                const double *B, const double *w, const double *amp, const double *damp, double *p0,
                double *p1, double *q0, double *q1, double *born, void *stream);

/*
 * Free functions of the reference module, on n planes of H x W:
 *   ss_lap_*                gpu.py:33-37 lap (zero on the two outer cells)
 *   ss_lap_adjoint_*        gpu.py:40-53 lap_adjoint (exact transpose of lap)
 *   ss_fold_edge_padding_*  gpu.py:56-67 fold_edge_padding: g (n, nz+2pad, nx+2pad) -> out (n, nz, nx)
 */
int ss_lap_f32(int64_t n, int H, int W, const float *in, float *out, void *stream);
int ss_lap_f64(int64_t n, int H, int W, const double *in, double *out, void *stream);
int ss_lap_adjoint_f32(int64_t n, int H, int W, const float *in, float *out, void *stream);
int ss_lap_adjoint_f64(int64_t n, int H, int W, const double *in, double *out, void *stream);
int ss_fold_edge_padding_f32(int64_t n, int nz, int nx, int pad, const float *g, float *out, void *stream);
int ss_fold_edge_padding_f64(int64_t n, int nz, int nx, int pad, const double *g, double *out, void *stream);

/*
 * Learned preconditioner (no reference counterpart: test-time refinement is
 * "not implemented yet" in the reference, README.md:25).  FP32, NCHW, 3x3
 * convolutions with zero "same" padding; w is (Cout, Cin, 3, 3), b is (Cout).
 *   ss_conv3x3_f32        y = conv(x, w) + b, optional ReLU      x (N,Cin,H,W) -> y (N,Cout,H,W)
 *   ss_conv3x3_dgrad_f32  dx = conv_transpose(dz, w)            dz (N,Cout,H,W) -> dx (N,Cin,H,W)
 *   ss_conv3x3_wgrad_f32  dw = sum dz x-shifted, db = sum dz   (deterministic reduction order)
 *   ss_relu_grad_f32      dz = dy * (y > 0)
 *   ss_mse_f32            grad = (p - t) / n ; partial[k] = block sums of (p - t)^2
 *   ss_adam_f32           Adam update with bias correction at step >= 1
 */
int ss_conv3x3_f32(int N, int Cin, int Cout, int H, int W, const float *x, const float *w, const float *b, float *y,
                   int relu, void *stream);
int ss_conv3x3_dgrad_f32(int N, int Cin, int Cout, int H, int W, const float *dz, const float *w, float *dx,
                         void *stream);
int ss_conv3x3_wgrad_f32(int N, int Cin, int Cout, int H, int W, const float *x, const float *dz, float *dw,
                         float *db, void *stream);
int ss_relu_grad_f32(int64_t n, const float *dy, const float *y, float *dz, void *stream);
int ss_mse_f32(int64_t n, const float *p, const float *t, float *grad, float *partial, int nblocks, void *stream);
int ss_adam_f32(int64_t n, float *p, const float *g, float *m, float *v, float lr, float beta1, float beta2,
                float eps, int step, void *stream);
/*
 * Refinement helpers (FP32):
 *   ss_precond_features_f32  x (B, 2, nz*nx) = (sum_s grad / max|.|, sum_s illum / max|.|), gscale (B) = max|sum_s grad|
 *                            from grad, illum (B, ns, nz*nx)
 *   ss_sumsq_diff_f32        out (B) = sum_i (a - c)^2 over each row of n values (data misfit per model, x2)
 *   ss_velocity_update_f32   out = clamp(v + alpha[b] * d, lo, hi) over B rows of n values
 */
int ss_precond_features_f32(int B, int ns, int nzx, const float *grad, const float *illum, float *x, float *gscale,
                            void *stream);
int ss_sumsq_diff_f32(int B, int64_t n, const float *a, const float *c, float *out, void *stream);
int ss_velocity_update_f32(int B, int64_t n, const float *v, const float *d, const float *alpha, float lo, float hi,
                           float *out, void *stream);

/*
 * First-break picks for the surface-velocity anchor, replacing the per-trace
 * loop of background.py:30-40 (amplitude.max over time and the first sample
 * above frac * peak).  For every trace (gather g, receiver r) of obs
 * (ngathers, nt, nx): peak[g, r] = max_t |obs|, onset[g, r] = first t with
 * |obs| > frac * peak (0 if none).  The threshold and comparisons are in double,
 * so picks on FP32 gathers equal NumPy's on the same data as float64.
 *   onset (ngathers, nx) int32, peak (ngathers, nx) in the element type
 */
int ss_first_breaks_f32(int64_t ngathers, int nt, int nx, const float *obs, double frac, int32_t *onset,
                        float *peak, void *stream);
int ss_first_breaks_f64(int64_t ngathers, int nt, int nx, const double *obs, double frac, int32_t *onset,
                        double *peak, void *stream);

/*
 * OpenFWI scores of B velocity maps (H x W), replacing physics/metrics.py:14-73
 * (norm, torch_ssim, metrics): out (B, 3) doubles = MAE and RMSE of the maps
 * normalised to [-1, 1] by (v - vmin) / (vmax - vmin) * 2 - 1, and SSIM of the
 * same maps shifted to [0, 1] with the 11 x 11 window (window, 121 doubles, the
 * outer product of the normalised sigma-1.5 Gaussian), zero padding and
 * c1 = 0.01^2, c2 = 0.03^2.  Computed in double; sums are tree reductions.
 */
int ss_velocity_scores_f32(int B, int H, int W, const float *pred, const float *truth, double vmin, double vmax,
                           const double *window, double *out, void *stream);
int ss_velocity_scores_f64(int B, int H, int W, const double *pred, const double *truth, double vmin,
                           double vmax, const double *window, double *out, void *stream);

/*
 * Mixed-size sweep (FP32).  Up to SS_SWEEP_MAX_GROUPS grid-size groups, each
 * a batch on its own padded grid with its own step count, shot count and
 * source columns, advance together: one tiled, temporally blocked launch per
 * step range covers every active (propagation, tile) item of every group,
 * and the launch sequence is replayed from a cached hipGraph.  A group whose
 * tiling is left at 0 and whose call covers the whole record from zero state
 * runs instead with its planes held on chip for the whole record: on the band
 * kernels (one cooperative launch; see ss_sweep_round_f32) for the 200 x 200
 * and 100 x 140 models with pad 40 and receivers on row 1, or on the
 * column-strip kernels (one workgroup per propagation; its state then stays in
 * buffer `phase`) for 70 x 70 with pad 40 and row 1.  Per cell the
 * arithmetic is the reference's, as in ss_forward_f32 / ss_adjoint_f32 over
 * [0, nt) with the whole Laplacian history:
 *   forward  gpu.py:111-115 and 126-135 (gathers, Laplacian history)
 *   adjoint  gpu.py:177-184 (imaging sum, illumination, source term)
 * The reference runs one 70 x 70 geometry per Sim (gpu.py:85, 95, 103, 126);
 * mixing sizes in one launch has no counterpart there.
 *
 * Per group (arrays as in ss_forward_f32 / ss_adjoint_f32, all FP32):
 *   geom              nb rows x ns shots on (nz+2 pad) x (nx+2 pad), nt steps
 *   tile_h, tile_w, depth   own-tile shape and steps per launch; 0 = the value
 *                     ss_sweep_autotune_f32 found for this grid and propagation count,
 *                     else a cost model
 *   c2 (nb,H,W), A, B (H,W), w (nt)
 *   amp               forward: source amplitude (nb, ns); adjoint: source derivative (nb, ns)
 *   res               adjoint: muted residual (nb, ns, nt, nx)
 *   state             scratch, 4 * nb * ns * H * W floats
 *   gathers           forward output (nb, ns, nt, nx) or NULL
 *   hist              Laplacian history (nb, ns, nt, H, W): forward writes, adjoint reads;
 *                     NULL for a forward that keeps no history (state and gathers only)
 *   gacc, illum, src_grad   adjoint outputs (nb,ns,H,W), (nb,ns,nz,nx), (nb,ns), overwritten
 *   t0, t1            step range of this call, [t0, t1) (t1 = 0: [0, nt)); hist then holds
 *                     (nb, ns, t1-t0, H, W) and gathers / res stay (nb, ns, nt, nx).  The
 *                     adjoint runs its range backwards.  A checkpointed gradient calls the
 *                     forward and the adjoint once per segment (gpu.py:169-184).
 *   resume            0: start from zero state and overwrite the adjoint outputs;
 *                     1: read the state from state buffer `phase` and add to the outputs
 *   phase             launch k of the call reads state buffer (k + phase) % 2 and writes the
 *                     other; after n launches the state is in buffer (n + phase) % 2.  The
 *                     state buffer b holds planes (p0, p1) or (m1, m2) at
 *                     state + (2 b + plane) * nb * ns * H * W.
 * Limits: H*W <= 131072, pad <= 44, nt*H*W < 2^29.
 * use_graph = 1 captures the launch sequence once per layout (group order,
 * tilings and pointers) and replays it; 0 launches directly.
 */
#define SS_SWEEP_MAX_GROUPS 4

typedef struct ss_sweep_group {
  ss_geom geom;
  int32_t tile_h, tile_w, depth;
  const float *c2, *A, *B, *w;
  const float *amp;
  const float *res;
  float *state;
  float *gathers;
  float *hist;
  float *gacc, *illum, *src_grad;
  int32_t t0, t1, resume, phase;
} ss_sweep_group;

int ss_sweep_forward_f32(int ngroups, const ss_sweep_group *groups, int use_graph, void *stream);
int ss_sweep_adjoint_f32(int ngroups, const ss_sweep_group *groups, int use_graph, void *stream);

/*
 * Autotuner: for each group, time up to max_trials candidate tilings (tile
 * shape and depth ranked by a halo-redundancy model; 0 = 12) as trial
 * launches of that group alone on its own buffers, write the fastest into
 * tile_h / tile_w / depth and remember it for the grid size and the group's
 * propagation count nb * ns (direction 0 = forward, 1 = adjoint).  Synchronises the stream; the trial launches leave
 * the group's outputs and state undefined, so run the sweep afterwards.
 */
int ss_sweep_autotune_f32(int ngroups, ss_sweep_group *groups, int direction, int max_trials, void *stream);

/*
 * Propagations per round of the band kernels on the current device, or 0 when
 * the grid has no band layout.  A group on a grid with a band layout (200 x 200
 * and 100 x 140 models, pad 40, receivers and sources on row 1) whose tiling is
 * left at 0 and whose call covers the whole record from zero state runs on the
 * band kernels: a team of workgroups, one per band of rows, holds each plane
 * for the whole record and trades the bands' edge rows every step, and one
 * cooperative launch runs the group's propagations in rounds of this many, so
 * batches of a multiple of it fill every round.  No counterpart in the
 * reference (one 70 x 70 geometry, gpu.py:85); it serves batch planning for
 * ss_sweep_forward_f32 / ss_sweep_adjoint_f32.
 */
int ss_sweep_round_f32(const ss_geom *geom);

/*
 * Count of band-kernel exchange waits on the current device that ran past their
 * bound (about a second of polls; the environment variable SHOTSWEEP_BAND_SPIN
 * sets the number of polls) since the library was loaded.  Such a wait fills the
 * band's halo with NaN and lets the launch finish, so the failure shows in the
 * outputs; with every workgroup of a team resident (a cooperative launch) it does
 * not occur.  Synchronous.  No counterpart in the reference (diagnostics of the
 * band kernels).
 */
int ss_sweep_band_errors(int *count);

#ifdef __cplusplus
}
#endif

#endif /* SHOTSWEEP_H */
