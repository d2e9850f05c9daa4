/*
 * ffm_soft.c -- FP64 restatement of the differentiable soft-splat mode:
 * the forward density and its adjoint (gradient of a loss with respect to
 * the transforms' affine coefficients and variation weights).
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker used by tests/ (and
 * nothing else); the product library (libffm.so) never links, loads or
 * calls it.
 *
 * The walk is the chaos game of ffm_oracle.c (shaders/core/main_template.wgsl:
 * selection :235-248 via orc_select, opacity draw :249, transform :355-384,
 * bad-value respawn :397-479, fuse :510-514) evaluated in double, with the
 * discrete choices (transform pick, opacity, julia's omega, julian's branch)
 * made from the same f32 draws as the device so both walk the same
 * branches. The reference has no soft splat (splat_size is hard-wired to
 * 1.0 and read by nothing, src/renderer/compute_kernel.rs:1354), so its
 * definition is this project's: each plotted sample adds density_scale x
 * bilinear weights around its continuous pixel position
 * (world_to_pixel, shaders/core/utilities.wgsl:417-445, without the floor,
 * minus half a pixel).
 *
 * Pinning: the adjoint is checked against central finite differences of
 * orc_soft_forward (tests), which pins both the Jacobians below and the
 * reverse sweep; the forward's total mass equals density_scale times the
 * number of plotted samples whose four taps are all in frame (tests); the
 * truncated adjoint equals the exact one when the window covers the whole
 * walk (tests).
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "ffm_oracle.h"

enum { SOFT_PLOT = 1, SOFT_RESPAWN = 2 };

/* V(x, y) and its Jacobian J = [dVx/dx, dVx/dy, dVy/dx, dVy/dy] in double.
 * `draw` says whether the variation consumes its random draw (the forward
 * draws only for a non-zero weight). Bodies: src/variations/defs/basic.rs
 * (linear :37, sinusoidal :100, spherical :130, swirl :162, horseshoe :198)
 * and advanced.rs (polar :32, disc :134, hyperbolic :213, julia :329,
 * julian :478, bubble :639, curl :1686). */
static int soft_var(int id, const float* p, double x, double y, orc_rng* rng, int draw,
                    double V[2], double J[4]) {
  const double PI = 3.14159265358979323846;
  double r2 = x * x + y * y;
  switch (id) {
    case FFM_VAR_LINEAR:
      V[0] = x; V[1] = y;
      J[0] = 1; J[1] = 0; J[2] = 0; J[3] = 1;
      return 0;
    case FFM_VAR_SINUSOIDAL:
      V[0] = sin(x); V[1] = sin(y);
      J[0] = cos(x); J[1] = 0; J[2] = 0; J[3] = cos(y);
      return 0;
    case FFM_VAR_SPHERICAL:
    case FFM_VAR_HYPERBOLIC: {
      double s = r2 + 1e-6;
      V[0] = x / s;
      J[0] = 1.0 / s - 2.0 * x * x / (s * s);
      J[1] = -2.0 * x * y / (s * s);
      if (id == FFM_VAR_SPHERICAL) {
        V[1] = y / s;
        J[2] = J[1];
        J[3] = 1.0 / s - 2.0 * y * y / (s * s);
      } else {
        V[1] = y;
        J[2] = 0;
        J[3] = 1;
      }
      return 0;
    }
    case FFM_VAR_SWIRL: {
      double S = sin(r2), C = cos(r2);
      V[0] = x * S - y * C;
      V[1] = x * C + y * S;
      J[0] = S + 2 * x * V[1]; J[1] = -C + 2 * y * V[1];
      J[2] = C - 2 * x * V[0]; J[3] = S - 2 * y * V[0];
      return 0;
    }
    case FFM_VAR_HORSESHOE: {
      double rho = sqrt(r2), r = rho + 1e-6;
      double ux = rho > 0 ? x / rho : 0, uy = rho > 0 ? y / rho : 0;
      V[0] = (x - y) * (x + y) / r;
      V[1] = 2 * x * y / r;
      J[0] = 2 * x / r - V[0] * ux / r; J[1] = -2 * y / r - V[0] * uy / r;
      J[2] = 2 * y / r - V[1] * ux / r; J[3] = 2 * x / r - V[1] * uy / r;
      return 0;
    }
    case FFM_VAR_POLAR:
    case FFM_VAR_DISC: {
      double th = atan2(x, y), rho = sqrt(r2);
      double dax = r2 > 0 ? y / (PI * r2) : 0, day = r2 > 0 ? -x / (PI * r2) : 0;
      double a = th / PI;
      if (id == FFM_VAR_POLAR) {
        V[0] = a; V[1] = rho - 1;
        J[0] = dax; J[1] = day;
        J[2] = rho > 0 ? x / rho : 0; J[3] = rho > 0 ? y / rho : 0;
      } else {
        double b = PI * rho, sb = sin(b), cb = cos(b);
        double dbx = rho > 0 ? PI * x / rho : 0, dby = rho > 0 ? PI * y / rho : 0;
        V[0] = a * sb; V[1] = a * cb;
        J[0] = dax * sb + a * cb * dbx; J[1] = day * sb + a * cb * dby;
        J[2] = dax * cb - a * sb * dbx; J[3] = day * cb - a * sb * dby;
      }
      return 0;
    }
    case FFM_VAR_JULIA: {
      double rho = sqrt(r2), th = atan2(y, x);
      double omega = (draw && orc_rng_nextf(rng) < 0.5f) ? PI : 0.0;
      double h = th / 2.0 + omega, sr = sqrt(rho);
      V[0] = sr * cos(h); V[1] = sr * sin(h);
      /* holomorphic: f' = V / (2 z) */
      double re = r2 > 0 ? (V[0] * x + V[1] * y) / (2 * r2) : 0;
      double im = r2 > 0 ? (V[1] * x - V[0] * y) / (2 * r2) : 0;
      J[0] = re; J[1] = -im; J[2] = im; J[3] = re;
      return 0;
    }
    case FFM_VAR_JULIAN: {
      float power = p[0], dist = p[1];
      double cp = (double)dist / fmax(fabs((double)power), 1e-30) / 2.0;
      double r = pow(r2, cp), th = atan2(y, x);
      /* the branch index is an f32 product on the device */
      float tr = draw ? floorf(fabsf(power) * orc_rng_nextf(rng)) : 0.0f;
      double t = (th + 2.0 * PI * (double)tr) / (double)power;
      double ct = cos(t), st = sin(t);
      V[0] = r * ct; V[1] = r * st;
      double drx = r2 > 0 ? cp * r * 2 * x / r2 : 0, dry = r2 > 0 ? cp * r * 2 * y / r2 : 0;
      double dtx = r2 > 0 ? -y / r2 / power : 0, dty = r2 > 0 ? x / r2 / power : 0;
      J[0] = drx * ct - r * st * dtx; J[1] = dry * ct - r * st * dty;
      J[2] = drx * st + r * ct * dtx; J[3] = dry * st + r * ct * dty;
      return 0;
    }
    case FFM_VAR_BUBBLE: {
      double s = r2 / 4.0 + 1.0;
      V[0] = x / s; V[1] = y / s;
      J[0] = 1 / s - x * x / (2 * s * s); J[1] = -x * y / (2 * s * s);
      J[2] = J[1]; J[3] = 1 / s - y * y / (2 * s * s);
      return 0;
    }
    case FFM_VAR_CURL: {
      double c1 = p[0], c2 = p[1];
      double re = 1.0 + c1 * x + c2 * (x * x - y * y), im = c1 * y + 2.0 * c2 * x * y;
      double d = re * re + im * im;
      V[0] = (x * re + y * im) / d;
      V[1] = (y * re - x * im) / d;
      /* f = z / w, f' = (1 - c2 z^2) / w^2 */
      double nr = 1.0 - c2 * (x * x - y * y), ni = -c2 * 2 * x * y;
      double wr = re * re - im * im, wi = 2 * re * im, wd = wr * wr + wi * wi;
      double fr = (nr * wr + ni * wi) / wd, fi = (ni * wr - nr * wi) / wd;
      J[0] = fr; J[1] = -fi; J[2] = fi; J[3] = fr;
      return 0;
    }
    default:
      V[0] = V[1] = 0;
      J[0] = J[1] = J[2] = J[3] = 0;
      return -1;
  }
}

int orc_soft_param_count(const ffm_flame_desc* f) {
  if (f->n_finals != 0 || f->n_xforms < 1 || f->n_xforms > FFM_SOFT_MAX_XFORMS ||
      f->n_vars < 1 || f->n_vars > FFM_SOFT_MAX_VARS)
    return -1;
  for (int k = 0; k < f->n_vars; k++)
    if (f->var_id[k] == FFM_VAR_WAVES) return -1;
  return f->n_xforms * (6 + f->n_vars);
}

typedef struct {
  const ffm_flame_desc* f;
  const double* th;  /* [n x 6 affine][n x v weights] */
  int n, v;
} soft_flame;

static const double* th_aff(const soft_flame* s, int i) { return s->th + 6 * i; }
static const double* th_w(const soft_flame* s, int i) { return s->th + 6 * s->n + s->v * i; }

/* One step of the walk; returns SOFT_* flags and the picked transform. */
static int soft_step(const soft_flame* s, uint32_t burn_in, orc_rng* rng, double* px,
                     double* py, uint32_t* fuse, int* xi_out) {
  const ffm_flame_desc* f = s->f;
  int xi = orc_select(f, orc_rng_nextf(rng));
  *xi_out = xi;
  const ffm_xform* xf = &f->xforms[xi];
  int plot = orc_rng_nextf(rng) < xf->opacity;
  const double* A = th_aff(s, xi);
  const double* W = th_w(s, xi);
  double x = *px, y = *py;
  double qx = A[0] * x + A[1] * y + A[4], qy = A[2] * x + A[3] * y + A[5];
  double rx = 0, ry = 0;
  for (int k = 0; k < s->v; k++) {
    if (W[k] == 0.0) continue;
    double V[2], J[4];
    soft_var(f->var_id[k], xf->var_param[k], qx, qy, rng, 1, V, J);
    rx += W[k] * V[0];
    ry += W[k] * V[1];
  }
  if (xf->post_enabled) {
    double tx = xf->post_a * rx + xf->post_b * ry + xf->post_e;
    double ty = xf->post_c * rx + xf->post_d * ry + xf->post_f;
    rx = tx;
    ry = ty;
  }
  if (!(fabs(rx) <= 1e32) || !(fabs(ry) <= 1e32)) {
    *px = orc_rng_nextf(rng) * 2.0 - 1.0;
    *py = orc_rng_nextf(rng) * 2.0 - 1.0;
    *fuse = burn_in;
    return SOFT_RESPAWN;
  }
  *px = rx;
  *py = ry;
  if (*fuse > 0) {
    (*fuse)--;
    return 0;
  }
  return plot ? SOFT_PLOT : 0;
}

typedef struct {
  double k;             /* d(pixel)/d(world) scale: zoom * min(W,H)/4 */
  double cr, sr, cx, cy;
  double pan_x, pan_y;
  int W, H;
} soft_view;

static soft_view make_soft_view(const ffm_view* v) {
  soft_view o;
  o.W = v->width;
  o.H = v->height;
  o.k = (double)v->zoom * (double)(v->width < v->height ? v->width : v->height) * 0.25;
  o.cr = cos((double)v->rotation);
  o.sr = sin((double)v->rotation);
  o.cx = v->width * 0.5 - 0.5;
  o.cy = v->height * 0.5 - 0.5;
  o.pan_x = v->pan_x;
  o.pan_y = v->pan_y;
  return o;
}

/* Continuous pixel position of a world point (pixel centres at integers). */
static int soft_pixel(const soft_view* o, double x, double y, double* sx, double* sy) {
  double tx = x - o->pan_x, ty = y - o->pan_y;
  *sx = o->cx + (tx * o->cr - ty * o->sr) * o->k;
  *sy = o->cy + (tx * o->sr + ty * o->cr) * o->k;
  return *sx > -1.0 && *sy > -1.0 && *sx < o->W && *sy < o->H;
}

static void walker_init(uint32_t tid, uint32_t seed, uint32_t burn_in, orc_rng* rng, double* x,
                        double* y, uint32_t* fuse) {
  orc_rng_init(rng, tid, seed);
  *x = orc_rng_nextf(rng) * 2.0 - 1.0;
  *y = orc_rng_nextf(rng) * 2.0 - 1.0;
  *fuse = burn_in;
}

typedef struct {
  const soft_flame* s;
  const ffm_soft_job* job;
  uint64_t w0, w1;
  double* density;        /* forward: W*H, this thread's */
  const double* g;        /* adjoint: dL/d density */
  double* grad;           /* adjoint: this thread's gradient */
} soft_task;

static void* forward_run(void* p) {
  soft_task* t = (soft_task*)p;
  const ffm_soft_job* j = t->job;
  soft_view o = make_soft_view(&j->view);
  for (uint64_t tid = t->w0; tid < t->w1; tid++) {
    orc_rng rng;
    double x, y;
    uint32_t fuse;
    walker_init((uint32_t)tid, j->seed, j->burn_in, &rng, &x, &y, &fuse);
    for (uint32_t it = 0; it < j->ipt; it++) {
      int xi;
      int fl = soft_step(t->s, j->burn_in, &rng, &x, &y, &fuse, &xi);
      if (!(fl & SOFT_PLOT)) continue;
      double sx, sy;
      if (!soft_pixel(&o, x, y, &sx, &sy)) continue;
      double x0 = floor(sx), y0 = floor(sy), ux = sx - x0, uy = sy - y0;
      int ix = (int)x0, iy = (int)y0;
      double wt[4] = {(1 - ux) * (1 - uy), ux * (1 - uy), (1 - ux) * uy, ux * uy};
      for (int c = 0; c < 4; c++) {
        int X = ix + (c & 1), Y = iy + (c >> 1);
        if (X >= 0 && X < o.W && Y >= 0 && Y < o.H)
          t->density[(size_t)Y * o.W + X] += j->density_scale * wt[c];
      }
    }
  }
  return NULL;
}

/* dL/d(world point) of one plotted sample: the bilinear gather of g at the
 * sample's position, differentiated, times density_scale. */
static void splat_grad(const soft_view* o, const double* g, double scale, double x, double y,
                       double lam[2]) {
  double sx, sy;
  if (!soft_pixel(o, x, y, &sx, &sy)) return;
  double x0 = floor(sx), y0 = floor(sy), ux = sx - x0, uy = sy - y0;
  int ix = (int)x0, iy = (int)y0;
  double G[4];
  for (int c = 0; c < 4; c++) {
    int X = ix + (c & 1), Y = iy + (c >> 1);
    G[c] = (X >= 0 && X < o->W && Y >= 0 && Y < o->H) ? g[(size_t)Y * o->W + X] : 0.0;
  }
  double gsx = (1 - uy) * (G[1] - G[0]) + uy * (G[3] - G[2]);
  double gsy = (1 - ux) * (G[2] - G[0]) + ux * (G[3] - G[1]);
  /* s = c + k R (p - pan): dL/dp = k R^T dL/ds */
  lam[0] += scale * o->k * (o->cr * gsx + o->sr * gsy);
  lam[1] += scale * o->k * (-o->sr * gsx + o->cr * gsy);
}

/* Pull lam (dL/d output point) back through step `xi` taken from the point
 * (x, y) with the RNG at `rs` before the step; adds the parameter gradient
 * to `grad` and leaves dL/d input point in lam. */
static void step_back(const soft_flame* s, double x, double y, orc_rng rs, int xi, double lam[2],
                      double* grad) {
  const ffm_flame_desc* f = s->f;
  const ffm_xform* xf = &f->xforms[xi];
  (void)orc_rng_nextf(&rs);  /* selection */
  (void)orc_rng_nextf(&rs);  /* opacity */
  const double* A = th_aff(s, xi);
  const double* W = th_w(s, xi);
  double qx = A[0] * x + A[1] * y + A[4], qy = A[2] * x + A[3] * y + A[5];
  double lv[2] = {lam[0], lam[1]};
  if (xf->post_enabled) {
    lv[0] = xf->post_a * lam[0] + xf->post_c * lam[1];
    lv[1] = xf->post_b * lam[0] + xf->post_d * lam[1];
  }
  double Js[4] = {0, 0, 0, 0};
  double* gw = grad + 6 * s->n + s->v * xi;
  for (int k = 0; k < s->v; k++) {
    double V[2], J[4];
    soft_var(f->var_id[k], xf->var_param[k], qx, qy, &rs, W[k] != 0.0, V, J);
    double gv = lv[0] * V[0] + lv[1] * V[1];
    if (isfinite(gv)) gw[k] += gv;
    if (W[k] != 0.0)
      for (int q = 0; q < 4; q++) Js[q] += W[k] * J[q];
  }
  double mx = Js[0] * lv[0] + Js[2] * lv[1];  /* mu = Js^T lv */
  double my = Js[1] * lv[0] + Js[3] * lv[1];
  double* ga = grad + 6 * xi;
  ga[0] += mx * x; ga[1] += mx * y; ga[4] += mx;
  ga[2] += my * x; ga[3] += my * y; ga[5] += my;
  lam[0] = A[0] * mx + A[2] * my;  /* A^T mu */
  lam[1] = A[1] * mx + A[3] * my;
}

static void* adjoint_run(void* p) {
  soft_task* t = (soft_task*)p;
  const ffm_soft_job* j = t->job;
  soft_view o = make_soft_view(&j->view);
  uint32_t n = j->ipt;
  double* px = malloc(sizeof(double) * (n + 1));
  double* py = malloc(sizeof(double) * (n + 1));
  orc_rng* rs = malloc(sizeof(orc_rng) * n);
  int* fl = malloc(sizeof(int) * n);
  int* xs = malloc(sizeof(int) * n);
  uint32_t K = j->truncate ? (j->window ? j->window : 1) : n;
  for (uint64_t tid = t->w0; tid < t->w1; tid++) {
    orc_rng rng;
    double x, y;
    uint32_t fuse;
    walker_init((uint32_t)tid, j->seed, j->burn_in, &rng, &x, &y, &fuse);
    for (uint32_t it = 0; it < n; it++) {
      px[it] = x;
      py[it] = y;
      rs[it] = rng;
      fl[it] = soft_step(t->s, j->burn_in, &rng, &x, &y, &fuse, &xs[it]);
    }
    px[n] = x;
    py[n] = y;
    double lam[2] = {0, 0};
    for (int64_t it = (int64_t)n - 1; it >= 0; it--) {
      if (fl[it] & SOFT_PLOT) splat_grad(&o, t->g, j->density_scale, px[it + 1], py[it + 1], lam);
      if (fl[it] & SOFT_RESPAWN) {
        lam[0] = lam[1] = 0;
      } else {
        step_back(t->s, px[it], py[it], rs[it], xs[it], lam, t->grad);
      }
      if ((uint32_t)it % K == 0) lam[0] = lam[1] = 0;  /* window start */
    }
  }
  free(px);
  free(py);
  free(rs);
  free(fl);
  free(xs);
  return NULL;
}

static int run_tasks(const soft_flame* s, const ffm_soft_job* job, int threads, int adjoint,
                     const double* g, double* out, size_t out_len) {
  if (threads < 1) threads = 1;
  if ((uint64_t)threads > job->walkers) threads = (int)(job->walkers ? job->walkers : 1);
  soft_task* tasks = calloc((size_t)threads, sizeof(soft_task));
  pthread_t* th = calloc((size_t)threads, sizeof(pthread_t));
  double* bufs = calloc((size_t)threads * out_len, sizeof(double));
  if (!tasks || !th || !bufs) {
    free(tasks); free(th); free(bufs);
    return -2;
  }
  for (int i = 0; i < threads; i++) {
    tasks[i].s = s;
    tasks[i].job = job;
    tasks[i].w0 = job->walkers * (uint64_t)i / (uint64_t)threads;
    tasks[i].w1 = job->walkers * (uint64_t)(i + 1) / (uint64_t)threads;
    tasks[i].g = g;
    if (adjoint) tasks[i].grad = bufs + (size_t)i * out_len;
    else tasks[i].density = bufs + (size_t)i * out_len;
    pthread_create(&th[i], NULL, adjoint ? adjoint_run : forward_run, &tasks[i]);
  }
  for (int i = 0; i < threads; i++) pthread_join(th[i], NULL);
  memset(out, 0, out_len * sizeof(double));
  for (int i = 0; i < threads; i++)
    for (size_t q = 0; q < out_len; q++) out[q] += bufs[(size_t)i * out_len + q];
  free(tasks);
  free(th);
  free(bufs);
  return 0;
}

int orc_soft_forward(const ffm_flame_desc* f, const double* theta, const ffm_soft_job* job,
                     double* density, int threads) {
  if (orc_soft_param_count(f) < 0) return -1;
  soft_flame s = {f, theta, f->n_xforms, f->n_vars};
  return run_tasks(&s, job, threads, 0, NULL, density,
                   (size_t)job->view.width * (size_t)job->view.height);
}

int orc_soft_adjoint(const ffm_flame_desc* f, const double* theta, const ffm_soft_job* job,
                     const double* grad_density, double* grad_theta, int threads) {
  int P = orc_soft_param_count(f);
  if (P < 0) return -1;
  soft_flame s = {f, theta, f->n_xforms, f->n_vars};
  return run_tasks(&s, job, threads, 1, grad_density, grad_theta, (size_t)P);
}

/* V and J of one variation at (x, y) in double, for tests of the device
 * Jacobians (no draw: julia's omega = 0, julian's branch 0). */
int orc_soft_variation(int id, const float* params, double x, double y, double* V, double* J) {
  return soft_var(id, params, x, y, NULL, 0, V, J);
}
