// extern "C" entry points of libshotsweep.so (declared in include/shotsweep.h).
// Validation happens here; kernels assume a checked geometry.
#include <cstdio>
#include <exception>
#include <string>

#include "ss_kernels.h"

namespace ss {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string &msg) { g_last_error = msg; }

int hip_status(hipError_t e, const char *what) {
  if (e == hipSuccess) return SS_OK;
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return SS_EHIP;
}

int check_geom(const ss_geom *g, Geom *out) {
  char buf[256];
  if (!g) {
    set_error("geometry is NULL");
    return SS_EINVAL;
  }
  if (g->nb < 1 || g->ns < 1 || g->ns > SS_MAX_SHOTS || g->nz < 1 || g->nx < 1 || g->pad < 2 || g->nt < 1) {
    snprintf(buf, sizeof buf, "invalid geometry nb=%d ns=%d nz=%d nx=%d pad=%d nt=%d (ns <= %d, pad >= 2)", g->nb,
             g->ns, g->nz, g->nx, g->pad, g->nt, SS_MAX_SHOTS);
    set_error(buf);
    return SS_EINVAL;
  }
  if (g->row < 0 || g->row >= g->nz) {
    snprintf(buf, sizeof buf, "source/receiver row %d outside 0..%d", g->row, g->nz - 1);
    set_error(buf);
    return SS_EINVAL;
  }
  const int64_t H = (int64_t)g->nz + 2 * g->pad, W = (int64_t)g->nx + 2 * g->pad;
  if (H * W > (1 << 30) || (int64_t)g->nt * g->nx > (1 << 30)) {
    set_error("grid too large");
    return SS_EINVAL;
  }
  out->nb = g->nb;
  out->ns = g->ns;
  out->nz = g->nz;
  out->nx = g->nx;
  out->pad = g->pad;
  out->nt = g->nt;
  out->row = g->row;
  out->H = (int)H;
  out->W = (int)W;
  for (int s = 0; s < SS_MAX_SHOTS; ++s) {
    out->src[s] = s < g->ns ? g->src[s] : 0;
    if (s < g->ns && (g->src[s] < 0 || g->src[s] >= g->nx)) {
      snprintf(buf, sizeof buf, "source column %d of shot %d outside 0..%d", g->src[s], s, g->nx - 1);
      set_error(buf);
      return SS_EINVAL;
    }
  }
  return SS_OK;
}

namespace {

int check_range(const Geom &g, int t0, int t1) {
  if (t0 < 0 || t1 > g.nt || t0 >= t1) {
    char buf[128];
    snprintf(buf, sizeof buf, "step range [%d, %d) outside [0, %d)", t0, t1, g.nt);
    set_error(buf);
    return SS_EINVAL;
  }
  return SS_OK;
}

int need(const void *p, const char *name) {
  if (p) return SS_OK;
  set_error(std::string(name) + " is NULL");
  return SS_EINVAL;
}

#define SS_TRY(expr)            \
  do {                          \
    int _rc = (expr);           \
    if (_rc != SS_OK) return _rc; \
  } while (0)

template <typename T>
int forward_impl(const ss_geom *gg, int t0, int t1, const T *c2, const T *A, const T *B, const T *w, const T *amp,
                 T *p0, T *p1, T *gathers, T *hist, T *ckpt, int ckpt_every, int impl, void *stream) {
  Geom g;
  SS_TRY(check_geom(gg, &g));
  SS_TRY(check_range(g, t0, t1));
  SS_TRY(need(c2, "c2"));
  SS_TRY(need(A, "A"));
  SS_TRY(need(B, "B"));
  SS_TRY(need(w, "w"));
  SS_TRY(need(amp, "amp"));
  SS_TRY(need(p0, "p0"));
  SS_TRY(need(p1, "p1"));
  if (ckpt && ckpt_every < 1) {
    set_error("ckpt_every must be >= 1");
    return SS_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  if constexpr (std::is_same<T, float>::value) {
    const bool fits = resident_fits(g);
    if (impl == SS_IMPL_RESIDENT && !fits) {
      set_error("resident kernels need the padded plane in LDS; this grid is too large");
      return SS_EUNSUP;
    }
    if (impl == SS_IMPL_RESIDENT || (impl == SS_IMPL_AUTO && fits))
      return forward_resident(g, t0, t1, c2, A, B, w, amp, p0, p1, gathers, hist, ckpt, ckpt_every, st);
  } else {
    if (impl == SS_IMPL_RESIDENT) {
      set_error("resident kernels are FP32 only");
      return SS_EUNSUP;
    }
  }
  return forward_stepwise<T>(g, t0, t1, c2, A, B, w, amp, p0, p1, gathers, hist, ckpt, ckpt_every, st);
}

template <typename T>
int adjoint_impl(const ss_geom *gg, int t0, int t1, const T *c2, const T *A, const T *B, const T *w, const T *sd,
                 const T *res, const T *hist, T *m1, T *m2, T *gacc, T *illum, T *src_grad, int impl,
                 void *stream) {
  Geom g;
  SS_TRY(check_geom(gg, &g));
  SS_TRY(check_range(g, t0, t1));
  SS_TRY(need(c2, "c2"));
  SS_TRY(need(A, "A"));
  SS_TRY(need(B, "B"));
  SS_TRY(need(w, "w"));
  SS_TRY(need(sd, "sd"));
  SS_TRY(need(res, "res"));
  SS_TRY(need(hist, "lap_hist"));
  SS_TRY(need(m1, "m1"));
  SS_TRY(need(m2, "m2"));
  SS_TRY(need(gacc, "gacc"));
  SS_TRY(need(illum, "illum"));
  SS_TRY(need(src_grad, "src_grad"));
  hipStream_t st = (hipStream_t)stream;
  if constexpr (std::is_same<T, float>::value) {
    const bool fits = resident_fits(g);
    if (impl == SS_IMPL_RESIDENT && !fits) {
      set_error("resident kernels need the padded plane in LDS; this grid is too large");
      return SS_EUNSUP;
    }
    if (impl == SS_IMPL_RESIDENT || (impl == SS_IMPL_AUTO && fits))
      return adjoint_resident(g, t0, t1, c2, A, B, w, sd, res, hist, m1, m2, gacc, illum, src_grad, st);
  } else {
    if (impl == SS_IMPL_RESIDENT) {
      set_error("resident kernels are FP32 only");
      return SS_EUNSUP;
    }
  }
  return adjoint_stepwise<T>(g, t0, t1, c2, A, B, w, sd, res, hist, m1, m2, gacc, illum, src_grad, st);
}

template <typename T>
int residual_impl(const ss_geom *gg, const T *sim, const T *obs, int obs_repeat, const T *v_mute, double dt,
                  double dx, double t_mute, double ramp, T *res, int weights_only, void *stream) {
  Geom g;
  SS_TRY(check_geom(gg, &g));
  SS_TRY(need(res, "res"));
  if (!weights_only) {
    SS_TRY(need(sim, "sim"));
    SS_TRY(need(obs, "obs"));
    if (obs_repeat < 1 || g.nb % obs_repeat != 0) {
      set_error("obs_repeat must be >= 1 and divide nb");
      return SS_EINVAL;
    }
  } else {
    SS_TRY(need(v_mute, "v_mute"));
  }
  if (!(dt > 0) || !(dx > 0) || !(ramp > 0)) {
    set_error("dt, dx and the ramp length must be positive");
    return SS_EINVAL;
  }
  return residual<T>(g, sim, obs, obs_repeat, v_mute, dt, dx, t_mute, ramp, res, weights_only, (hipStream_t)stream);
}

template <typename T>
int gradient_impl(const ss_geom *gg, const T *gacc, const T *v, const T *src_grad, T k, T *grad, void *stream) {
  Geom g;
  SS_TRY(check_geom(gg, &g));
  SS_TRY(need(gacc, "gacc"));
  SS_TRY(need(v, "v"));
  SS_TRY(need(src_grad, "src_grad"));
  SS_TRY(need(grad, "grad"));
  return gradient_epilogue<T>(g, gacc, v, src_grad, k, grad, (hipStream_t)stream);
}

template <typename T>
int image_impl(const ss_geom *gg, const T *grad, const T *illum, T eps, int laplacian, T *img, void *stream) {
  Geom g;
  SS_TRY(check_geom(gg, &g));
  SS_TRY(need(grad, "grad"));
  SS_TRY(need(illum, "illum"));
  SS_TRY(need(img, "img"));
  if (!(eps > 0)) {
    set_error("illumination regularization eps must be positive");
    return SS_EINVAL;
  }
  return image_epilogue<T>(g, grad, illum, eps, laplacian, img, (hipStream_t)stream);
}

template <typename T>
int born_impl(const ss_geom *gg, int t0, int t1, const T *c2, const T *dc2, const T *A, const T *B, const T *w,
              const T *amp, const T *damp, T *p0, T *p1, T *q0, T *q1, T *born, void *stream) {
  Geom g;
  SS_TRY(check_geom(gg, &g));
  SS_TRY(check_range(g, t0, t1));
  SS_TRY(need(c2, "c2"));
  SS_TRY(need(dc2, "dc2"));
  SS_TRY(need(A, "A"));
  SS_TRY(need(B, "B"));
  SS_TRY(need(w, "w"));
  SS_TRY(need(amp, "amp"));
  SS_TRY(need(damp, "damp"));
  SS_TRY(need(p0, "p0"));
  SS_TRY(need(p1, "p1"));
  SS_TRY(need(q0, "q0"));
  SS_TRY(need(q1, "q1"));
  return born_stepwise<T>(g, t0, t1, c2, dc2, A, B, w, amp, damp, p0, p1, q0, q1, born, (hipStream_t)stream);
}

template <typename T>
int lap_impl(int64_t n, int H, int W, const T *in, T *out, int adjoint, void *stream) {
  if (n < 0 || H < 1 || W < 1) {
    set_error("lap: invalid shape");
    return SS_EINVAL;
  }
  if (n == 0) return SS_OK;
  SS_TRY(need(in, "in"));
  SS_TRY(need(out, "out"));
  return lap_op<T>(n, H, W, in, out, adjoint, (hipStream_t)stream);
}

template <typename T>
int fold_impl(int64_t n, int nz, int nx, int pad, const T *g, T *out, void *stream) {
  if (n < 0 || nz < 1 || nx < 1 || pad < 0) {
    set_error("fold_edge_padding: invalid shape");
    return SS_EINVAL;
  }
  if (n == 0) return SS_OK;
  SS_TRY(need(g, "g"));
  SS_TRY(need(out, "out"));
  return fold_op<T>(n, nz, nx, pad, g, out, (hipStream_t)stream);
}

}  // namespace
}  // namespace ss

namespace ss {
namespace {

template <typename T>
int first_breaks_impl(int64_t ngathers, int nt, int nx, const T *obs, double frac, int32_t *onset, T *peak,
                      void *stream) {
  if (ngathers < 1 || nt < 1 || nx < 1 || !obs || !onset || !peak) {
    set_error("first_breaks: invalid arguments");
    return SS_EINVAL;
  }
  if (!(frac > 0.0 && frac < 1.0)) {
    set_error("onset threshold frac must lie between zero and one");
    return SS_EINVAL;
  }
  return first_breaks<T>(ngathers, nt, nx, obs, frac, onset, peak, (hipStream_t)stream);
}

template <typename T>
int scores_impl(int B, int H, int W, const T *pred, const T *truth, double vmin, double vmax, const double *window,
                double *out, void *stream) {
  if (B < 1 || H < 1 || W < 1 || !pred || !truth || !window || !out || !(vmax > vmin)) {
    set_error("velocity_scores: invalid arguments");
    return SS_EINVAL;
  }
  return velocity_scores<T>(B, H, W, pred, truth, vmin, vmax, window, out, (hipStream_t)stream);
}

}  // namespace
}  // namespace ss

using namespace ss;

extern "C" {

const char *ss_last_error(void) { return ss::g_last_error.c_str(); }

int ss_version(void) { return 8; }

int64_t ss_resident_max_cells(void) { return resident_max_cells(); }

int64_t ss_workspace_bytes(const ss_geom *gp, int segment, int elem_size) {
  return rtm_workspace_bytes(gp, segment, elem_size);
}

int ss_forward_f32(const ss_geom *g, int t0, int t1, const float *c2, const float *A, const float *B,
                   const float *w, const float *amp, float *p0, float *p1, float *gathers, float *lap_hist,
                   float *ckpt, int ckpt_every, int impl, void *stream) {
  return forward_impl<float>(g, t0, t1, c2, A, B, w, amp, p0, p1, gathers, lap_hist, ckpt, ckpt_every, impl,
                             stream);
}
int ss_forward_f64(const ss_geom *g, int t0, int t1, const double *c2, const double *A, const double *B,
                   const double *w, const double *amp, double *p0, double *p1, double *gathers, double *lap_hist,
                   double *ckpt, int ckpt_every, int impl, void *stream) {
  return forward_impl<double>(g, t0, t1, c2, A, B, w, amp, p0, p1, gathers, lap_hist, ckpt, ckpt_every, impl,
                              stream);
}

int ss_adjoint_f32(const ss_geom *g, int t0, int t1, const float *c2, const float *A, const float *B,
                   const float *w, const float *sd, const float *res, const float *lap_hist, float *m1, float *m2,
                   float *gacc, float *illum, float *src_grad, int impl, void *stream) {
  return adjoint_impl<float>(g, t0, t1, c2, A, B, w, sd, res, lap_hist, m1, m2, gacc, illum, src_grad, impl,
                             stream);
}
int ss_adjoint_f64(const ss_geom *g, int t0, int t1, const double *c2, const double *A, const double *B,
                   const double *w, const double *sd, const double *res, const double *lap_hist, double *m1,
                   double *m2, double *gacc, double *illum, double *src_grad, int impl, void *stream) {
  return adjoint_impl<double>(g, t0, t1, c2, A, B, w, sd, res, lap_hist, m1, m2, gacc, illum, src_grad, impl,
                              stream);
}

int ss_residual_f32(const ss_geom *g, const float *sim, const float *obs, int obs_repeat, const float *v_mute,
                    double dt_s, double dx_m, double t_mute, double ramp, float *res, void *stream) {
  return residual_impl<float>(g, sim, obs, obs_repeat, v_mute, dt_s, dx_m, t_mute, ramp, res, 0, stream);
}
int ss_residual_f64(const ss_geom *g, const double *sim, const double *obs, int obs_repeat, const double *v_mute,
                    double dt_s, double dx_m, double t_mute, double ramp, double *res, void *stream) {
  return residual_impl<double>(g, sim, obs, obs_repeat, v_mute, dt_s, dx_m, t_mute, ramp, res, 0, stream);
}
int ss_mute_weights_f32(const ss_geom *g, const float *v_mute, double dt_s, double dx_m, double t_mute, double ramp,
                        float *out, void *stream) {
  return residual_impl<float>(g, nullptr, nullptr, 1, v_mute, dt_s, dx_m, t_mute, ramp, out, 1, stream);
}
int ss_mute_weights_f64(const ss_geom *g, const double *v_mute, double dt_s, double dx_m, double t_mute,
                        double ramp, double *out, void *stream) {
  return residual_impl<double>(g, nullptr, nullptr, 1, v_mute, dt_s, dx_m, t_mute, ramp, out, 1, stream);
}

int ss_gradient_f32(const ss_geom *g, const float *gacc, const float *v, const float *src_grad, float k, float *grad,
                    void *stream) {
  return gradient_impl<float>(g, gacc, v, src_grad, k, grad, stream);
}
int ss_gradient_f64(const ss_geom *g, const double *gacc, const double *v, const double *src_grad, double k,
                    double *grad, void *stream) {
  return gradient_impl<double>(g, gacc, v, src_grad, k, grad, stream);
}

int ss_image_f32(const ss_geom *g, const float *grad, const float *illum, float eps, int laplacian, float *img,
                 void *stream) {
  return image_impl<float>(g, grad, illum, eps, laplacian, img, stream);
}
int ss_image_f64(const ss_geom *g, const double *grad, const double *illum, double eps, int laplacian, double *img,
                 void *stream) {
  return image_impl<double>(g, grad, illum, eps, laplacian, img, stream);
}

int ss_born_f32(const ss_geom *g, int t0, int t1, const float *c2, const float *dc2, const float *A, const float *B,
                const float *w, const float *amp, const float *damp, float *p0, float *p1, float *q0, float *q1,
                float *born, void *stream) {
  return born_impl<float>(g, t0, t1, c2, dc2, A, B, w, amp, damp, p0, p1, q0, q1, born, stream);
}
int ss_born_f64(const ss_geom *g, int t0, int t1, const double *c2, const double *dc2, const double *A,
                const double *B, const double *w, const double *amp, const double *damp, double *p0, double *p1,
                double *q0, double *q1, double *born, void *stream) {
  return born_impl<double>(g, t0, t1, c2, dc2, A, B, w, amp, damp, p0, p1, q0, q1, born, stream);
}

int ss_lap_f32(int64_t n, int H, int W, const float *in, float *out, void *stream) {
  return lap_impl<float>(n, H, W, in, out, 0, stream);
}
int ss_lap_f64(int64_t n, int H, int W, const double *in, double *out, void *stream) {
  return lap_impl<double>(n, H, W, in, out, 0, stream);
}
int ss_lap_adjoint_f32(int64_t n, int H, int W, const float *in, float *out, void *stream) {
  return lap_impl<float>(n, H, W, in, out, 1, stream);
}
int ss_lap_adjoint_f64(int64_t n, int H, int W, const double *in, double *out, void *stream) {
  return lap_impl<double>(n, H, W, in, out, 1, stream);
}
int ss_fold_edge_padding_f32(int64_t n, int nz, int nx, int pad, const float *g, float *out, void *stream) {
  return fold_impl<float>(n, nz, nx, pad, g, out, stream);
}
int ss_fold_edge_padding_f64(int64_t n, int nz, int nx, int pad, const double *g, double *out, void *stream) {
  return fold_impl<double>(n, nz, nx, pad, g, out, stream);
}

static int conv_shape_ok(int N, int Cin, int Cout, int H, int W) {
  if (N < 1 || Cin < 1 || Cout < 1 || H < 1 || W < 1 || Cin > 32 || Cout > 1024) {
    set_error("conv3x3: invalid shape (1 <= Cin <= 32)");
    return 0;
  }
  return 1;
}

int ss_conv3x3_f32(int N, int Cin, int Cout, int H, int W, const float *x, const float *w, const float *b, float *y,
                   int relu, void *stream) {
  if (!conv_shape_ok(N, Cin, Cout, H, W)) return SS_EINVAL;
  if (!x || !w || !y) {
    set_error("conv3x3: NULL pointer");
    return SS_EINVAL;
  }
  return conv3x3(N, Cin, Cout, H, W, x, w, b, y, relu, 0, (hipStream_t)stream);
}

int ss_conv3x3_dgrad_f32(int N, int Cin, int Cout, int H, int W, const float *dz, const float *w, float *dx,
                         void *stream) {
  if (!conv_shape_ok(N, Cout, Cin, H, W)) return SS_EINVAL;
  if (!dz || !w || !dx) {
    set_error("conv3x3_dgrad: NULL pointer");
    return SS_EINVAL;
  }
  return conv3x3(N, Cout, Cin, H, W, dz, w, nullptr, dx, 0, 1, (hipStream_t)stream);
}

int ss_conv3x3_wgrad_f32(int N, int Cin, int Cout, int H, int W, const float *x, const float *dz, float *dw,
                         float *db, void *stream) {
  if (!conv_shape_ok(N, Cin, Cout, H, W)) return SS_EINVAL;
  if (!x || !dz || !dw) {
    set_error("conv3x3_wgrad: NULL pointer");
    return SS_EINVAL;
  }
  return conv3x3_wgrad(N, Cin, Cout, H, W, x, dz, dw, db, (hipStream_t)stream);
}

int ss_relu_grad_f32(int64_t n, const float *dy, const float *y, float *dz, void *stream) {
  if (n < 0 || (n && (!dy || !y || !dz))) {
    set_error("relu_grad: invalid arguments");
    return SS_EINVAL;
  }
  return n ? relu_mask(n, dy, y, dz, (hipStream_t)stream) : SS_OK;
}

int ss_mse_f32(int64_t n, const float *p, const float *t, float *grad, float *partial, int nblocks, void *stream) {
  if (n < 1 || nblocks < 1 || nblocks > 65536 || !p || !t || !grad || !partial) {
    set_error("mse: invalid arguments");
    return SS_EINVAL;
  }
  return mse(n, p, t, grad, partial, nblocks, (hipStream_t)stream);
}

int ss_adam_f32(int64_t n, float *p, const float *g, float *m, float *v, float lr, float beta1, float beta2,
                float eps, int step, void *stream) {
  if (n < 1 || step < 1 || !p || !g || !m || !v) {
    set_error("adam: invalid arguments");
    return SS_EINVAL;
  }
  return adam(n, p, g, m, v, lr, beta1, beta2, eps, step, (hipStream_t)stream);
}

int ss_precond_features_f32(int B, int ns, int nzx, const float *grad, const float *illum, float *x, float *gscale,
                            void *stream) {
  if (B < 1 || ns < 1 || nzx < 1 || !grad || !illum || !x || !gscale) {
    set_error("precond_features: invalid arguments");
    return SS_EINVAL;
  }
  return precond_features(B, ns, nzx, grad, illum, x, gscale, (hipStream_t)stream);
}

int ss_sumsq_diff_f32(int B, int64_t n, const float *a, const float *c, float *out, void *stream) {
  if (B < 1 || n < 1 || !a || !c || !out) {
    set_error("sumsq_diff: invalid arguments");
    return SS_EINVAL;
  }
  return sumsq_diff(B, n, a, c, out, (hipStream_t)stream);
}

int ss_velocity_update_f32(int B, int64_t n, const float *v, const float *d, const float *alpha, float lo, float hi,
                           float *out, void *stream) {
  if (B < 1 || n < 1 || !v || !d || !alpha || !out || !(lo <= hi)) {
    set_error("velocity_update: invalid arguments");
    return SS_EINVAL;
  }
  return velocity_update(B, n, v, d, alpha, lo, hi, out, (hipStream_t)stream);
}

int ss_first_breaks_f32(int64_t ngathers, int nt, int nx, const float *obs, double frac, int32_t *onset,
                        float *peak, void *stream) {
  return first_breaks_impl<float>(ngathers, nt, nx, obs, frac, onset, peak, stream);
}

int ss_first_breaks_f64(int64_t ngathers, int nt, int nx, const double *obs, double frac, int32_t *onset,
                        double *peak, void *stream) {
  return first_breaks_impl<double>(ngathers, nt, nx, obs, frac, onset, peak, stream);
}

int ss_velocity_scores_f32(int B, int H, int W, const float *pred, const float *truth, double vmin, double vmax,
                           const double *window, double *out, void *stream) {
  return scores_impl<float>(B, H, W, pred, truth, vmin, vmax, window, out, stream);
}

int ss_velocity_scores_f64(int B, int H, int W, const double *pred, const double *truth, double vmin,
                           double vmax, const double *window, double *out, void *stream) {
  return scores_impl<double>(B, H, W, pred, truth, vmin, vmax, window, out, stream);
}

int ss_sweep_forward_f32(int ngroups, const ss_sweep_group *groups, int use_graph, void *stream) {
  try {
    return sweep_run(ngroups, groups, 0, use_graph, (hipStream_t)stream);
  } catch (const std::exception &e) {
    set_error(std::string("sweep forward: ") + e.what());
    return SS_EHIP;
  }
}

int ss_sweep_adjoint_f32(int ngroups, const ss_sweep_group *groups, int use_graph, void *stream) {
  try {
    return sweep_run(ngroups, groups, 1, use_graph, (hipStream_t)stream);
  } catch (const std::exception &e) {
    set_error(std::string("sweep adjoint: ") + e.what());
    return SS_EHIP;
  }
}

int ss_sweep_autotune_f32(int ngroups, ss_sweep_group *groups, int direction, int max_trials, void *stream) {
  if (direction != 0 && direction != 1) {
    set_error("sweep autotune: direction must be 0 (forward) or 1 (adjoint)");
    return SS_EINVAL;
  }
  try {
    return sweep_autotune(ngroups, groups, direction, max_trials, (hipStream_t)stream);
  } catch (const std::exception &e) {
    set_error(std::string("sweep autotune: ") + e.what());
    return SS_EHIP;
  }
}

int ss_sweep_band_errors(int *count) {
  if (!count) {
    set_error("sweep band errors: count is required");
    return SS_EINVAL;
  }
  try {
    return band_errors(count);
  } catch (const std::exception &e) {
    set_error(std::string("sweep band errors: ") + e.what());
    return SS_EHIP;
  }
}

int ss_sweep_round_f32(const ss_geom *geom) {
  Geom g;
  const int rc = check_geom(geom, &g);
  if (rc != SS_OK) return rc;
  try {
    return band_round(g);
  } catch (const std::exception &e) {
    set_error(std::string("sweep round: ") + e.what());
    return SS_EHIP;
  }
}

}  // extern "C"
