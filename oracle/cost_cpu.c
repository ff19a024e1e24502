/*
 * FP64 CPU restatement of SMILI's imaging cost, C + OpenMP.
 *
 * TEST INFRASTRUCTURE / CPU BASELINE ONLY: built by oracle/Makefile and
 * used by tests and by bench.py's cpu_baseline leg; the product path never
 * links it.
 *
 * It follows the structure of the reference's hot path,
 * fftim2d.calc_cost (src/fortlib/fftim2d.f90:387-607):
 *   calc_chisq (src/fortlib/fftlib.f90:250-382)
 *     - I1d_I2d_fwd + one type-2 NUFFT (FINUFFT_fwd, wrapfinufft.f90:94-128)
 *     - chi-square loops for fcv / amp / ca / cp (fftlib.f90:384-655) with
 *       per-thread residual arrays summed afterwards (the OpenMP
 *       REDUCTION(+:Vresre,Vresim) of the reference)
 *     - NUFFT_adj_resid (fftlib.f90:152-174): two type-1 NUFFTs, of Vresre
 *       and of Vresim, then Re(.) + Im(.)
 *   calc_cost_reg (src/fortlib/image.f90:12-200): l1 / kl / gs / tv / tsv
 *     (tfd / cen / sm are not restated here; the numpy oracle has them)
 * The NUFFT is the exponential-of-semicircle method FINUFFT implements
 * (Barnett, Magland & af Klinteberg 2019): width 7, upsampling 2,
 * beta = 2.30 * 7, own mixed-radix (2, 3, 4, 5) FFT.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/smili_hip.h"

#define W 7
#define ZEROEPS 1e-10
static const double BETA = 2.30 * W;

typedef double complex cplx;

typedef struct {
  int n, nst, radix[16];
  cplx *tw; /* exp(+2 pi i t / n) */
} fftplan;

typedef struct smili_cpu {
  int nx, ny, nf1, nf2, npix, m;
  int *pix; /* window pixel -> iy*nx+ix */
  double *u, *v;
  double *cx, *cy;
  fftplan p1, p2;
  smili_reg reg;
  double *l1w, *tvw, *tsvw, *klw, *gsw;
  /* data (copied) */
  int nfcv, namp, ncp, nca;
  int *ifcv, *iamp, *icp, *ica;
  double *fre, *fim, *fvar, *amp, *avar, *cp, *cvar, *ca, *qvar;
  double wfcv, wamp, wcp, wca;
  int nthreads;
} smili_cpu;

/* ------------------------------------------------------------------------- */
static double es_phi(double z) {
  if (fabs(z) > 1.0) return 0.0;
  return exp(-BETA * z * z / (1.0 + sqrt(1.0 - z * z)));
}

static void correction(int n, int nf, double *c) {
  /* c(k) = 1 / ((w/2) int_{-1}^{1} phi(z) cos(pi k w z / nf) dz), midpoint rule
     on 4000 nodes (phi is smooth and vanishes at +-1) */
  const int Q = 4000;
  for (int i = 0; i < n; ++i) {
    const double k = i - n / 2;
    double s = 0.0;
    for (int q = 0; q < Q; ++q) {
      const double z = -1.0 + (q + 0.5) * (2.0 / Q);
      s += es_phi(z) * cos(M_PI * k * W * z / nf);
    }
    s *= 2.0 / Q;
    c[i] = 1.0 / (0.5 * W * s);
  }
}

static int next_smooth(int n) {
  for (;; ++n) {
    int m = n;
    while (m % 2 == 0) m /= 2;
    while (m % 3 == 0) m /= 3;
    while (m % 5 == 0) m /= 5;
    if (m == 1) return n;
  }
}

static void plan_init(fftplan *p, int n) {
  p->n = n;
  p->nst = 0;
  int r = n;
  while (r % 4 == 0) { p->radix[p->nst++] = 4; r /= 4; }
  while (r % 2 == 0) { p->radix[p->nst++] = 2; r /= 2; }
  while (r % 3 == 0) { p->radix[p->nst++] = 3; r /= 3; }
  while (r % 5 == 0) { p->radix[p->nst++] = 5; r /= 5; }
  p->tw = malloc(sizeof(cplx) * n);
  for (int t = 0; t < n; ++t) p->tw[t] = cexp(I * 2.0 * M_PI * t / n);
}

/* Stockham autosort FFT with the +i sign; a and s have n entries (stride 1).
   Returns the buffer holding the result. */
static cplx *fft1d(const fftplan *p, cplx *a, cplx *s) {
  const int n = p->n;
  int Ns = 1;
  for (int st = 0; st < p->nst; ++st) {
    const int r = p->radix[st];
    const int nbf = n / r;
    const int twstep = n / (Ns * r);
    for (int i = 0; i < nbf; ++i) {
      const int k = i % Ns;
      cplx v[5];
      for (int q = 0; q < r; ++q) v[q] = a[i + q * nbf];
      if (Ns > 1)
        for (int q = 1; q < r; ++q) v[q] *= p->tw[q * k * twstep];
      cplx y[5];
      if (r == 2) {
        y[0] = v[0] + v[1];
        y[1] = v[0] - v[1];
      } else if (r == 4) {
        const cplx t0 = v[0] + v[2], t1 = v[0] - v[2], t2 = v[1] + v[3], t3 = I * (v[1] - v[3]);
        y[0] = t0 + t2; y[2] = t0 - t2; y[1] = t1 + t3; y[3] = t1 - t3;
      } else {
        for (int mm = 0; mm < r; ++mm) {
          cplx acc = 0;
          for (int q = 0; q < r; ++q) acc += v[q] * p->tw[((q * mm) % r) * (n / r)];
          y[mm] = acc;
        }
      }
      const int base = (i / Ns) * Ns * r + k;
      for (int q = 0; q < r; ++q) s[base + q * Ns] = y[q];
    }
    cplx *t = a; a = s; s = t;
    Ns *= r;
  }
  return a;
}

/* 2-D FFT of g [nf2][nf1] in place (rows, then columns) */
static void fft2d(smili_cpu *P, cplx *g) {
  const int n1 = P->nf1, n2 = P->nf2;
#pragma omp parallel
  {
    cplx *a = malloc(sizeof(cplx) * (n1 > n2 ? n1 : n2));
    cplx *b = malloc(sizeof(cplx) * (n1 > n2 ? n1 : n2));
#pragma omp for schedule(static)
    for (int r = 0; r < n2; ++r) {
      memcpy(a, g + (long)r * n1, sizeof(cplx) * n1);
      cplx *o = fft1d(&P->p1, a, b);
      memcpy(g + (long)r * n1, o, sizeof(cplx) * n1);
    }
#pragma omp for schedule(static)
    for (int c = 0; c < n1; ++c) {
      for (int r = 0; r < n2; ++r) a[r] = g[(long)r * n1 + c];
      cplx *o = fft1d(&P->p2, a, b);
      for (int r = 0; r < n2; ++r) g[(long)r * n1 + c] = o[r];
    }
    free(a);
    free(b);
  }
}

static void taps(double x, int nf, int *l0, double *w) {
  double t = fmod(x * nf / (2.0 * M_PI), (double)nf);
  if (t < 0) t += nf;
  const double l = ceil(t - 0.5 * W);
  const double d = t - l;
  for (int q = 0; q < W; ++q) w[q] = es_phi((d - q) * (2.0 / W));
  int li = ((int)l) % nf;
  if (li < 0) li += nf;
  *l0 = li;
}

/* type 2: V_j = sum_k f_k exp(+i k.x_j) (image I2d [ny][nx]) */
static void nufft2(smili_cpu *P, const double *I2d, cplx *V, cplx *g) {
  const int nx = P->nx, ny = P->ny, n1 = P->nf1, n2 = P->nf2;
  memset(g, 0, sizeof(cplx) * (long)n1 * n2);
  for (int iy = 0; iy < ny; ++iy) {
    int r = iy - ny / 2;
    if (r < 0) r += n2;
    for (int ix = 0; ix < nx; ++ix) {
      int c = ix - nx / 2;
      if (c < 0) c += n1;
      g[(long)r * n1 + c] = I2d[(long)iy * nx + ix] * P->cx[ix] * P->cy[iy];
    }
  }
  fft2d(P, g);
#pragma omp parallel for schedule(static)
  for (int j = 0; j < P->m; ++j) {
    int lx, ly;
    double wx[W], wy[W];
    taps(P->u[j], n1, &lx, wx);
    taps(P->v[j], n2, &ly, wy);
    cplx acc = 0;
    for (int a = 0; a < W; ++a) {
      const int row = (ly + a) % n2;
      cplx ra = 0;
      for (int b = 0; b < W; ++b) ra += g[(long)row * n1 + (lx + b) % n1] * wx[b];
      acc += ra * wy[a];
    }
    V[j] = acc;
  }
}

/* type 1: F_k = sum_j c_j exp(+i k.x_j) -> out [ny][nx] complex */
static void nufft1(smili_cpu *P, const cplx *c, cplx *out, cplx *g, cplx *priv, int nth) {
  const int nx = P->nx, ny = P->ny, n1 = P->nf1, n2 = P->nf2;
  const long G = (long)n1 * n2;
#pragma omp parallel
  {
    int t = 0;
#ifdef _OPENMP
    t = omp_get_thread_num();
#endif
    cplx *mine = priv + (long)t * G;
    memset(mine, 0, sizeof(cplx) * G);
#pragma omp for schedule(static)
    for (int j = 0; j < P->m; ++j) {
      int lx, ly;
      double wx[W], wy[W];
      taps(P->u[j], n1, &lx, wx);
      taps(P->v[j], n2, &ly, wy);
      for (int a = 0; a < W; ++a) {
        const int row = (ly + a) % n2;
        const cplx ca = c[j] * wy[a];
        for (int b = 0; b < W; ++b) mine[(long)row * n1 + (lx + b) % n1] += ca * wx[b];
      }
    }
#pragma omp for schedule(static)
    for (long i = 0; i < G; ++i) {
      cplx s = 0;
      for (int k = 0; k < nth; ++k) s += priv[(long)k * G + i];
      g[i] = s;
    }
  }
  fft2d(P, g);
  for (int iy = 0; iy < ny; ++iy) {
    int r = iy - ny / 2;
    if (r < 0) r += n2;
    for (int ix = 0; ix < nx; ++ix) {
      int cc = ix - nx / 2;
      if (cc < 0) cc += n1;
      out[(long)iy * nx + ix] = g[(long)r * n1 + cc] * P->cx[ix] * P->cy[iy];
    }
  }
}

/* ------------------------------------------------------------------------- */
static double *dup(const double *a, long n) {
  if (!a || n <= 0) return NULL;
  double *b = malloc(sizeof(double) * n);
  memcpy(b, a, sizeof(double) * n);
  return b;
}

static int *dupi(const int32_t *a, long n) {
  if (!a || n <= 0) return NULL;
  int *b = malloc(sizeof(int) * n);
  for (long i = 0; i < n; ++i) b[i] = a[i];
  return b;
}

int smili_cpu_create(const smili_desc *d, smili_cpu **out) {
  if (!d || !out) return -1;
  smili_cpu *P = calloc(1, sizeof(smili_cpu));
  P->nx = d->nx; P->ny = d->ny; P->npix = d->npix; P->m = d->nuv;
  P->nf1 = next_smooth(2 * d->nx > 2 * W ? 2 * d->nx : 2 * W);
  P->nf2 = next_smooth(2 * d->ny > 2 * W ? 2 * d->ny : 2 * W);
  P->pix = malloc(sizeof(int) * d->npix);
  for (int i = 0; i < d->npix; ++i) P->pix[i] = (d->yidx[i] - 1) * d->nx + (d->xidx[i] - 1);
  P->u = dup(d->u, d->nuv);
  P->v = dup(d->v, d->nuv);
  P->cx = malloc(sizeof(double) * d->nx);
  P->cy = malloc(sizeof(double) * d->ny);
  correction(d->nx, P->nf1, P->cx);
  correction(d->ny, P->nf2, P->cy);
  plan_init(&P->p1, P->nf1);
  plan_init(&P->p2, P->nf2);
  P->reg = d->reg;
  P->l1w = d->reg.l1_l > 0 ? dup(d->reg.l1_wgt, d->npix) : NULL;
  P->tvw = d->reg.tv_l > 0 ? dup(d->reg.tv_wgt, d->npix) : NULL;
  P->tsvw = d->reg.tsv_l > 0 ? dup(d->reg.tsv_wgt, d->npix) : NULL;
  P->klw = d->reg.kl_l > 0 ? dup(d->reg.kl_wgt, d->npix) : NULL;
  P->gsw = d->reg.gs_l > 0 ? dup(d->reg.gs_wgt, d->npix) : NULL;
  const smili_data *D = &d->data;
  P->nfcv = D->nfcv; P->namp = D->namp; P->ncp = D->ncp; P->nca = D->nca;
  P->ifcv = dupi(D->uvidxfcv, D->nfcv);
  P->iamp = dupi(D->uvidxamp, D->namp);
  P->icp = dupi(D->uvidxcp, 3L * D->ncp);
  P->ica = dupi(D->uvidxca, 4L * D->nca);
  P->fre = dup(D->vfcv_re, D->nfcv); P->fim = dup(D->vfcv_im, D->nfcv); P->fvar = dup(D->varfcv, D->nfcv);
  P->amp = dup(D->vamp, D->namp); P->avar = dup(D->varamp, D->namp);
  P->cp = dup(D->cp, D->ncp); P->cvar = dup(D->varcp, D->ncp);
  P->ca = dup(D->ca, D->nca); P->qvar = dup(D->varca, D->nca);
  P->wfcv = D->wfcv; P->wamp = D->wamp; P->wcp = D->wcp; P->wca = D->wca;
  P->nthreads = 1;
#ifdef _OPENMP
  P->nthreads = omp_get_max_threads();
#endif
  *out = P;
  return 0;
}

void smili_cpu_destroy(smili_cpu *P) {
  if (!P) return;
  void *ptrs[] = {P->pix, P->u, P->v, P->cx, P->cy, P->p1.tw, P->p2.tw, P->l1w, P->tvw, P->tsvw, P->klw,
                  P->gsw, P->ifcv, P->iamp, P->icp, P->ica, P->fre, P->fim, P->fvar, P->amp, P->avar,
                  P->cp, P->cvar, P->ca, P->qvar};
  for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); ++i) free(ptrs[i]);
  free(P);
}

int smili_cpu_threads(smili_cpu *P) { return P ? P->nthreads : 0; }

static double smabs(double x) { return sqrt(x * x + ZEROEPS); }

int smili_cpu_cost_grad(smili_cpu *P, const double *x, double *cost, double *grad, double *terms) {
  const int nx = P->nx, ny = P->ny, M = P->m, nth = P->nthreads;
  const long G = (long)P->nf1 * P->nf2;
  double *I2d = calloc((size_t)nx * ny, sizeof(double));
  for (int i = 0; i < P->npix; ++i) I2d[P->pix[i]] = x[i];
  cplx *g = malloc(sizeof(cplx) * G);
  cplx *priv = malloc(sizeof(cplx) * G * nth);
  cplx *V = malloc(sizeof(cplx) * M);
  nufft2(P, I2d, V, g);
  /* chi-square loops with per-thread residual arrays */
  double *rr = calloc((size_t)M * nth, sizeof(double));
  double *ri = calloc((size_t)M * nth, sizeof(double));
  double cf = 0, cam = 0, ccp = 0, cca = 0;
#pragma omp parallel reduction(+ : cf, cam, ccp, cca)
  {
    int t = 0;
#ifdef _OPENMP
    t = omp_get_thread_num();
#endif
    double *Rr = rr + (long)t * M, *Ri = ri + (long)t * M;
#pragma omp for schedule(static) nowait
    for (int i = 0; i < P->nfcv; ++i) {
      const int j = abs(P->ifcv[i]) - 1;
      const double s = P->ifcv[i] > 0 ? 1.0 : -1.0;
      const cplx Vm = s > 0 ? V[j] : conj(V[j]);
      const cplx r = (P->fre[i] + I * P->fim[i]) - Vm;
      cf += (creal(r) * creal(r) + cimag(r) * cimag(r)) / P->fvar[i] * P->wfcv;
      const double f = -2.0 / P->fvar[i] * P->wfcv;
      Rr[j] += f * creal(r);
      Ri[j] += f * cimag(r) * s;
    }
#pragma omp for schedule(static) nowait
    for (int i = 0; i < P->namp; ++i) {
      const int j = abs(P->iamp[i]) - 1;
      const double mod = cabs(V[j]);
      const double r = P->amp[i] - mod;
      cam += r * r / P->avar[i] * P->wamp;
      const double f = -2.0 * r / P->avar[i] / mod * P->wamp;
      Rr[j] += f * creal(V[j]);
      Ri[j] += f * cimag(V[j]);
    }
#pragma omp for schedule(static) nowait
    for (int i = 0; i < P->nca; ++i) {
      int j[4];
      double a[4];
      for (int k = 0; k < 4; ++k) {
        j[k] = abs(P->ica[4L * i + k]) - 1;
        a[k] = cabs(V[j[k]]);
      }
      const double model = log(a[0]) + log(a[1]) - log(a[2]) - log(a[3]);
      const double r = P->ca[i] - model;
      cca += r * r / P->qvar[i] * P->wca;
      const double f = -2.0 * r / P->qvar[i] * P->wca;
      for (int k = 0; k < 4; ++k) {
        const double q = (k < 2 ? f : -f) / (a[k] * a[k]);
        Rr[j[k]] += q * creal(V[j[k]]);
        Ri[j[k]] += q * cimag(V[j[k]]);
      }
    }
#pragma omp for schedule(static) nowait
    for (int i = 0; i < P->ncp; ++i) {
      int j[3];
      double s[3], model = 0;
      for (int k = 0; k < 3; ++k) {
        j[k] = abs(P->icp[3L * i + k]) - 1;
        s[k] = P->icp[3L * i + k] > 0 ? 1.0 : -1.0;
        model += carg(V[j[k]]) * s[k];
      }
      const double r = P->cp[i] - model;
      ccp += 2.0 * (1.0 - cos(r)) / P->cvar[i] * P->wcp;
      const double f = -2.0 * sin(r) / P->cvar[i] * P->wcp;
      for (int k = 0; k < 3; ++k) {
        const double a2 = creal(V[j[k]]) * creal(V[j[k]]) + cimag(V[j[k]]) * cimag(V[j[k]]);
        Rr[j[k]] -= f / a2 * cimag(V[j[k]]) * s[k];
        Ri[j[k]] += f / a2 * creal(V[j[k]]) * s[k];
      }
    }
  }
  cplx *cre = malloc(sizeof(cplx) * M), *cim = malloc(sizeof(cplx) * M);
#pragma omp parallel for schedule(static)
  for (int j = 0; j < M; ++j) {
    double a = 0, b = 0;
    for (int t = 0; t < nth; ++t) {
      a += rr[(long)t * M + j];
      b += ri[(long)t * M + j];
    }
    cre[j] = a;
    cim[j] = b;
  }
  /* NUFFT_adj_resid: two type-1 transforms, Re(A Vresre) + Im(A Vresim) */
  cplx *F1 = malloc(sizeof(cplx) * nx * ny), *F2 = malloc(sizeof(cplx) * nx * ny);
  nufft1(P, cre, F1, g, priv, nth);
  nufft1(P, cim, F2, g, priv, nth);
  /* regularizers (l1, kl, gs, tv, tsv) + gradient */
  const smili_reg *R = &P->reg;
  double l1 = 0, tv = 0, tsv = 0, kl = 0, gs = 0;
#pragma omp parallel for schedule(static) reduction(+ : l1, tv, tsv, kl, gs)
  for (int k = 0; k < P->npix; ++k) {
    const int p = P->pix[k], ix = p % nx, iy = p / nx;
    const double xi = x[k];
    double gr = creal(F1[p]) + cimag(F2[p]);
    double gg = 0;
    if (R->l1_l > 0) {
      l1 += R->l1_l * P->l1w[k] * smabs(xi);
      gg += R->l1_l * P->l1w[k] * xi / smabs(xi);
    }
    if (R->kl_l > 0) {
      const double a = smabs(xi), Pp = P->klw[k];
      kl += R->kl_l * (a * log(a / Pp) + Pp / M_E);
      gg += R->kl_l * xi / a * (log(a / Pp) + 1);
    }
    if (R->gs_l > 0) {
      const double a = smabs(xi), Pp = P->gsw[k];
      gs += R->gs_l * (a * (log(a / Pp) - 1) + Pp);
      gg += R->gs_l * xi / a * log(a / Pp);
    }
    if (R->tv_l > 0 || R->tsv_l > 0) {
      const double c0 = I2d[p];
      const double dx = ix + 1 < nx ? I2d[p + 1] - c0 : 0, dy = iy + 1 < ny ? I2d[p + nx] - c0 : 0;
      if (R->tv_l > 0) {
        double t = 0, e = sqrt(dx * dx + dy * dy);
        if (e > ZEROEPS) t -= (dx + dy) / e;
        if (ix > 0) {
          const double ax = c0 - I2d[p - 1];
          const double ay = iy + 1 < ny ? I2d[p - 1 + nx] - I2d[p - 1] : 0;
          e = sqrt(ax * ax + ay * ay);
          t = e > ZEROEPS ? t + ax / e : 0;
        }
        if (iy > 0) {
          const double ay = c0 - I2d[p - nx];
          const double ax = ix + 1 < nx ? I2d[p - nx + 1] - I2d[p - nx] : 0;
          e = sqrt(ax * ax + ay * ay);
          t = e > ZEROEPS ? t + ay / e : 0;
        }
        tv += R->tv_l * P->tvw[k] * sqrt(dx * dx + dy * dy);
        gg += R->tv_l * P->tvw[k] * t;
      }
      if (R->tsv_l > 0) {
        double t = 0;
        if (ix + 1 < nx) t -= 2 * (I2d[p + 1] - c0);
        if (iy + 1 < ny) t -= 2 * (I2d[p + nx] - c0);
        if (ix > 0) t += 2 * (c0 - I2d[p - 1]);
        if (iy > 0) t += 2 * (c0 - I2d[p - nx]);
        tsv += R->tsv_l * P->tsvw[k] * (dx * dx + dy * dy);
        gg += R->tsv_l * P->tsvw[k] * t;
      }
    }
    grad[k] = gr + gg;
  }
  double chisq = 0;
  if (P->nfcv) chisq += cf;
  if (P->namp) chisq += cam;
  if (P->nca) chisq += cca;
  if (P->ncp) chisq += ccp;
  const double reg = l1 + tv + tsv + kl + gs;
  if (terms) {
    double T[SMILI_NTERMS] = {chisq,
                              P->nfcv ? cf / P->wfcv / (2.0 * P->nfcv) : 0,
                              P->namp ? cam / P->wamp / P->namp : 0,
                              P->ncp ? ccp / P->wcp / P->ncp : 0,
                              P->nca ? cca / P->wca / P->nca : 0,
                              reg, l1, tv, tsv, kl, gs, 0, 0, 0, chisq + reg};
    memcpy(terms, T, sizeof(T));
  }
  if (cost) *cost = chisq + reg;
  free(I2d); free(g); free(priv); free(V); free(rr); free(ri); free(cre); free(cim); free(F1); free(F2);
  return 0;
}
