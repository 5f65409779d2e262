// Internal declarations shared by the kernel translation units and the ABI.
#pragma once

#include "ss_common.h"

namespace ss {

template <typename T>
int forward_stepwise(const Geom &g, int t0, int t1, const T *c2, const T *A, const T *B, const T *w, const T *amp,
                     T *p0, T *p1, T *gathers, T *hist, T *ckpt, int ckpt_every, hipStream_t st);

template <typename T>
int adjoint_stepwise(const Geom &g, int t0, int t1, const T *c2, const T *A, const T *B, const T *w, const T *sd,
                     const T *res, const T *hist, T *m1, T *m2, T *gacc, T *illum, T *src_grad, hipStream_t st);

template <typename T>
int born_stepwise(const Geom &g, int t0, int t1, const T *c2, const T *dc2, const T *A, const T *B, const T *w,
                  const T *amp, const T *damp, T *p0, T *p1, T *q0, T *q1, T *born, hipStream_t st);

template <typename T>
int residual(const Geom &g, const T *sim, const T *obs, int obs_repeat, const T *v_mute, double dt, double dx,
             double t_mute, double ramp, T *res, int weights_only, hipStream_t st);

template <typename T>
int gradient_epilogue(const Geom &g, const T *gacc, const T *v, const T *src_grad, T k, T *grad, hipStream_t st);

template <typename T>
int image_epilogue(const Geom &g, const T *grad, const T *illum, T eps, int laplacian, T *img, hipStream_t st);

template <typename T>
int lap_op(int64_t n, int H, int W, const T *in, T *out, int adjoint, hipStream_t st);

template <typename T>
int fold_op(int64_t n, int nz, int nx, int pad, const T *g, T *out, hipStream_t st);

// Learned-preconditioner kernels (FP32).
int conv3x3(int N, int Cin, int Cout, int H, int W, const float *x, const float *w, const float *b, float *y,
            int relu, int flip, hipStream_t st);
int conv3x3_wgrad(int N, int Cin, int Cout, int H, int W, const float *x, const float *dz, float *dw, float *db,
                  hipStream_t st);
int relu_mask(int64_t n, const float *dy, const float *y, float *dz, hipStream_t st);
int mse(int64_t n, const float *p, const float *t, float *grad, float *partial, int nblocks, hipStream_t st);
int adam(int64_t n, float *p, const float *g, float *m, float *v, float lr, float b1, float b2, float eps, int step,
         hipStream_t st);

int precond_features(int B, int ns, int nzx, const float *grad, const float *illum, float *x, float *gscale,
                     hipStream_t st);
int sumsq_diff(int B, int64_t n, const float *a, const float *c, float *out, hipStream_t st);
int velocity_update(int B, int64_t n, const float *v, const float *d, const float *alpha, float lo, float hi,
                    float *out, hipStream_t st);

// Resident FP32 kernels (plane held in LDS, one workgroup per propagation).
int64_t resident_max_cells();
bool resident_fits(const Geom &g);
// exact_rim = false: the 150 x 150 plane's history is left unspecified on the two-cell
// rim (only for the native driver, whose adjoint_resident with the c2 * A plane never
// uses it); every other output is unchanged.
int forward_resident(const Geom &g, int t0, int t1, const float *c2, const float *A, const float *B, const float *w,
                     const float *amp, float *p0, float *p1, float *gathers, float *hist, float *ckpt,
                     int ckpt_every, hipStream_t st, bool exact_rim = true);
// First-break picks (onset index, peak |g|) per trace of (ngathers, nt, nx) gathers.
template <typename T>
int first_breaks(int64_t ngathers, int nt, int nx, const T *obs, double frac, int *onset, T *peak, hipStream_t st);

// OpenFWI scores (MAE, RMSE, SSIM) of B velocity maps of H x W cells; window is 11 x 11 doubles.
template <typename T>
int velocity_scores(int B, int H, int W, const T *pred, const T *truth, double vmin, double vmax,
                    const double *window, double *out, hipStream_t st);
size_t scores_lds_bytes(int H, int W);

// Native gradient / images driver (ss_rtm.hip).
int64_t rtm_workspace_bytes(const ss_geom *g, int segment, int elem_size);

// Mixed-size tiled sweep (FP32); dir 0 = forward, 1 = adjoint.
int sweep_run(int ngroups, const ss_sweep_group *groups, int dir, int use_graph, hipStream_t st);
int sweep_autotune(int ngroups, ss_sweep_group *groups, int dir, int max_trials, hipStream_t st);
bool sweep_tuned(int dir, int H, int W, int pad, int64_t nprop);  // tiling tuned for this grid and propagation count
int sweep_launches(const ss_sweep_group &group, int dir, hipStream_t st);  // launches of one call on st (< 0: error)
bool sweep_fits(const Geom &g);                             // within the tiled kernels' limits

// Band kernels (ss_band.hip): planes too large for one workgroup held by a team of
// workgroups, one band of rows each, that trade their boundary rows every step
// (cooperative launch).  band_fits: a band layout exists for the grid and direction;
// band_run runs a sweep group's whole step range in one launch; band_errors reads the
// count of exchange waits that ran past spin_limit polls (their bands' halos became NaN).
bool band_fits(const Geom &g, int dir);
int band_run(const ss_sweep_group &grp, const Geom &g, int dir, int spin_limit, hipStream_t st);
int band_errors(int *count);
int band_round(const Geom &g);  // propagations per round of a band launch (0: no band layout)

// c2a: NULL, or the plane c2 * A of c2a_plane (used by the 150 x 150 paired kernel,
// see resident_uses_c2a; other shapes ignore it).
int adjoint_resident(const Geom &g, int t0, int t1, const float *c2, const float *A, const float *B, const float *w,
                     const float *sd, const float *res, const float *hist, float *m1, float *m2, float *gacc,
                     float *illum, float *src_grad, hipStream_t st, const float *c2a = nullptr);
bool resident_uses_c2a(const Geom &g);
// out (nb, H, W) = c2 * A on the interior cells, 0 on the two-cell rim
int c2a_plane(int64_t nb, int H, int W, const float *c2, const float *A, float *out, hipStream_t st);

}  // namespace ss
