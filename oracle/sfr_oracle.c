/*
 * ORACLE -- test infrastructure only.
 *
 * CPU restatement of the reference's linear-solve path, used as the checker
 * of the GPU implementation and as the timed CPU baseline.  Only the tests,
 * the smoke check and the benchmark's CPU-baseline leg may load it; the
 * product library never does.
 *
 * Followed line by line for operation order (serial dot is one left-to-right
 * accumulator, so results are reproducible bit for bit):
 *   dot              reference src/HccaBlas.c:393-427
 *   matVecCsr        reference src/HccaBlas.c:3244-3268
 *   matVecCsrSym     reference src/HccaBlas.c:3297-3331
 *   matVecCsrD       reference src/HccaBlas.c:3358-3382
 *   matVecCsrDSym    reference src/HccaBlas.c:2620-2651
 *   matVecCsrC       reference src/HccaBlas.c:4223-4260
 *   matVecEllPack    reference src/HccaBlas.c:4700-4752 (widths 6/4/3 only;
 *                    other widths leave y untouched, as the reference does)
 *   setMatVec        reference src/Solv.c:732-919 (serial branch)
 *   preCondDiag      reference src/PreCond.c:18-26
 *   pcg              reference src/Cg.c:47-209, driven as src/Solv.c:93-145
 *   pbicgstab        reference src/Bicgstab.c:50-245, as src/Solv.c:150-212
 *   pbicgstabl2      reference src/Bicgstab.c:1172-1378, as src/Solv.c:217-295
 *   gmres            reference src/Iters.c:587-790, as src/Solv.c:302-372
 *   pminres          reference src/Iters.c:309-555, as src/Solv.c:380-445
 *   symOrtho         reference src/Iters.c:1171-1208
 * The OpenMP variants (orc_*_omp) keep the same recurrences with parallel
 * loops and reductions over full rows; they are the all-cores CPU baseline
 * (the reference's pcgOmp, src/Cg.c:490-715, plays that role there).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_CSR 1
#define ORC_CSRD 2
#define ORC_CSRC 3
#define ORC_ELLPACK 4
#define ORC_CSRDCOO 5
#define ORC_CSRCCOO 6

typedef struct orc_stats {
  int status;          /* 0 converged, 1 maxit */
  unsigned iters;
  double xKx, norm_x, norm_r, norm_r_m, norm_b, conv, time_s;
} orc_stats;

typedef void (*matvec_fn)(int, const int *, const int *, const double *,
                          const double *, const double *, double *);

static double now(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

double orc_dot(const double *x1, const double *x2, int n)
{
  double tmp = 0.0;
  for (int i = 0; i < n; i++) tmp += x1[i] * x2[i];
  return tmp;
}

static void mv_csr(int n, const int *ia, const int *ja, const double *a,
                   const double *ad, const double *x, double *y)
{
  (void)ad;
  for (int i = 0; i < n; i++) {
    double tmp = 0.0;
    for (int j = ia[i]; j < ia[i + 1]; j++) tmp += a[j] * x[ja[j]];
    y[i] = tmp;
  }
}

static void mv_csr_sym(int n, const int *ia, const int *ja, const double *a,
                       const double *ad, const double *x, double *y)
{
  (void)ad;
  for (int i = 0; i < n; i++) {
    const double xi = x[i];
    double tmp = y[i] = 0.0;
    for (int j = ia[i]; j < ia[i + 1]; j++) {
      const int jak = ja[j];
      const double s = a[j];
      if (i == jak) tmp += s * xi;
      else { tmp += s * x[jak]; y[jak] += s * xi; }
    }
    y[i] = tmp;
  }
}

static void mv_csrd(int n, const int *ia, const int *ja, const double *a,
                    const double *ad, const double *x, double *y)
{
  for (int i = 0; i < n; i++) {
    double tmp = ad[i] * x[i];
    for (int j = ia[i]; j < ia[i + 1]; j++) tmp += a[j] * x[ja[j]];
    y[i] = tmp;
  }
}

static void mv_csrd_sym(int n, const int *ia, const int *ja, const double *al,
                        const double *ad, const double *x, double *y)
{
  y[0] = ad[0] * x[0];
  for (int i = 1; i < n; i++) {
    const double xi = x[i];
    double tmp = ad[i] * xi;
    for (int k = ia[i]; k < ia[i + 1]; k++) {
      const int jak = ja[k];
      const double s = al[k];
      tmp += s * x[jak];
      y[jak] += s * xi;
    }
    y[i] = tmp;
  }
}

static void mv_csrc(int n, const int *ia, const int *ja, const double *a,
                    const double *ad, const double *x, double *y)
{
  const int nad = ia[n] - ia[0];
  const double *au = a + nad, *al = a;
  y[0] = ad[0] * x[0];
  for (int i = 1; i < n; i++) {
    const double xi = x[i];
    double tmp = ad[i] * xi;
    for (int k = ia[i]; k < ia[i + 1]; k++) {
      const int jak = ja[k];
      tmp += al[k] * x[jak];
      y[jak] += au[k] * xi;
    }
    y[i] = tmp;
  }
}

static void mv_ell(int n, const int *ia, const int *ja, const double *a,
                   const double *ad, const double *x, double *y)
{
  const int s = ia[0];
  if (s != 6 && s != 4 && s != 3) return;
  for (int i = 0; i < n; i++) {
    double t = ad[i] * x[i];
    for (int f = 0; f < s; f++) t += a[i * s + f] * x[ja[i * s + f]];
    y[i] = t;
  }
}

static matvec_fn pick(int storage, int unsym)
{
  switch (storage) {
  case ORC_CSR: return unsym ? mv_csr : mv_csr_sym;
  case ORC_CSRD: case ORC_CSRDCOO: return unsym ? mv_csrd : mv_csrd_sym;
  case ORC_CSRC: case ORC_CSRCCOO: return unsym ? mv_csrc : mv_csrd_sym;
  case ORC_ELLPACK: return mv_ell;
  default: return NULL;
  }
}

/* y = A x with the storage's reference matvec (for checks) */
int orc_matvec(int storage, int unsym, int n, const int *ia, const int *ja,
               const double *al, const double *ad, const double *x, double *y)
{
  matvec_fn mv = pick(storage, unsym);
  if (!mv) return -1;
  mv(n, ia, ja, al, ad, x, y);
  return 0;
}

int orc_pcg(int n, int nad, const int *ia, const int *ja, const double *al,
            const double *ad, const double *b, double *x, double tol,
            unsigned maxit, int newX, int storage, int unsym, orc_stats *st)
{
  (void)nad;
  matvec_fn matvec = pick(storage, unsym);
  if (!matvec || n <= 0) return -1;
  double *m = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n);
  double *r = malloc(sizeof(double) * n), *p = malloc(sizeof(double) * n);
  if (!m || !z || !r || !p) { free(m); free(z); free(r); free(p); return -2; }
  for (int i = 0; i < n; i++) m[i] = 1.0 / ad[i];          /* preCondDiag */
  memset(z, 0, sizeof(double) * n);
  const double t0 = now();
  unsigned j;
  if (newX) for (int i = 0; i < n; i++) x[i] = 0.0;
  for (int i = 0; i < n; i++) z[i] = b[i] * m[i];
  double d = orc_dot(b, z, n);
  const double norm_b = sqrt(fabs(d));
  const double conv = tol * norm_b;
  matvec(n, ia, ja, al, ad, x, z);
  for (int i = 0; i < n; i++) {
    r[i] = b[i] - z[i];
    z[i] = r[i] * m[i];
    p[i] = z[i];
  }
  d = orc_dot(r, z, n);
  for (j = 0; j < maxit; j++) {
    matvec(n, ia, ja, al, ad, p, z);
    const double alpha = d / orc_dot(z, p, n);
    for (int i = 0; i < n; i++) {
      x[i] += alpha * p[i];
      r[i] -= alpha * z[i];
      z[i] = r[i] * m[i];
    }
    const double di = orc_dot(r, z, n);
    const double beta = di / d;
    for (int i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
    d = di;
    if (sqrt(fabs(d)) < conv) break;
  }
  matvec(n, ia, ja, al, ad, x, z);
  const double xKx = orc_dot(x, z, n);
  const double norm = sqrt(orc_dot(x, x, n));
  for (int i = 0; i < n; i++) {
    r[i] = b[i] - z[i];
    z[i] = r[i] * m[i];
  }
  const double norm_r_m = sqrt(fabs(orc_dot(r, z, n)));
  const double norm_r = sqrt(orc_dot(r, r, n));
  const double t1 = now();
  if (st) {
    st->status = (j == maxit) ? 1 : 0;
    st->iters = j + 1;
    st->xKx = xKx; st->norm_x = norm; st->norm_r = norm_r;
    st->norm_r_m = norm_r_m; st->norm_b = norm_b; st->conv = conv;
    st->time_s = t1 - t0;
  }
  free(m); free(z); free(r); free(p);
  return 0;
}

int orc_pbicgstab(int n, int nad, const int *ia, const int *ja,
                  const double *al, const double *ad, const double *b,
                  double *x, double tol, unsigned maxit, int newX,
                  int storage, int unsym, orc_stats *st)
{
  (void)nad;
  matvec_fn matvec = pick(storage, unsym);
  if (!matvec || n <= 0) return -1;
  double *buf = calloc((size_t)n * 8, sizeof(double));
  if (!buf) return -2;
  double *m = buf, *z = buf + n, *r = buf + 2 * (size_t)n, *t = buf + 3 * (size_t)n;
  double *v = buf + 4 * (size_t)n, *p = buf + 5 * (size_t)n, *r0 = buf + 6 * (size_t)n;
  for (int i = 0; i < n; i++) m[i] = 1.0 / ad[i];
  const double t0 = now();
  unsigned j;
  if (newX) for (int i = 0; i < n; i++) x[i] = 0.0;
  double d = orc_dot(b, b, n);
  const double norm_b = sqrt(d);
  const double conv = tol * norm_b;
  matvec(n, ia, ja, al, ad, x, z);
  for (int i = 0; i < n; i++) {
    r0[i] = b[i] - z[i];
    p[i] = r0[i];
    r[i] = r0[i];
    z[i] = p[i] * m[i];
  }
  for (j = 0; j < maxit; j++) {
    matvec(n, ia, ja, al, ad, z, v);
    const double rr0 = orc_dot(r, r0, n);
    const double alpha = rr0 / orc_dot(v, r0, n);
    for (int i = 0; i < n; i++) {
      x[i] += alpha * z[i];
      r[i] -= alpha * v[i];
      z[i] = r[i] * m[i];
    }
    d = orc_dot(r, r, n);
    if (sqrt(d) < conv) break;
    matvec(n, ia, ja, al, ad, z, t);
    const double w = orc_dot(t, r, n) / orc_dot(t, t, n);
    for (int i = 0; i < n; i++) {
      x[i] += w * z[i];
      r[i] -= w * t[i];
    }
    d = orc_dot(r, r, n);
    if (sqrt(d) < conv) break;
    const double beta = (orc_dot(r, r0, n) / rr0) * (alpha / w);
    for (int i = 0; i < n; i++) {
      p[i] = r[i] + beta * (p[i] - w * v[i]);
      z[i] = p[i] * m[i];
    }
  }
  matvec(n, ia, ja, al, ad, x, z);
  const double xKx = orc_dot(x, z, n);
  const double norm = sqrt(orc_dot(x, x, n));
  for (int i = 0; i < n; i++) r[i] = b[i] - z[i];
  const double norm_r = sqrt(orc_dot(r, r, n));
  double nrm = 0.0;
  for (int i = 0; i < n; i++) nrm += r[i] * (r[i] * m[i]);
  const double t1 = now();
  if (st) {
    st->status = (j == maxit) ? 1 : 0;
    st->iters = j + 1;
    st->xKx = xKx; st->norm_x = norm; st->norm_r = norm_r;
    st->norm_r_m = sqrt(fabs(nrm)); st->norm_b = norm_b; st->conv = conv;
    st->time_s = t1 - t0;
  }
  free(buf);
  return 0;
}

/* GMRES(k) with left Jacobi preconditioning, reference src/Iters.c:587-790,
 * driven as src/Solv.c:302-372 (k = nKrylov = 15, cycles = maxIt).
 * Kept as the reference runs it:
 *   - g_0 = D^-1 (b - A x); Arnoldi on D^-1 A with modified Gram-Schmidt;
 *   - the in-cycle convergence test is an empty statement (Iters.c:702), so
 *     every cycle runs all k Arnoldi steps;
 *   - the back substitution for y stops its inner sum at j < k-1
 *     (Iters.c:714-719), so y_{k-1} never enters the rows above it, and the
 *     update x += sum_{j<k-1} y_j g_j leaves out the last basis vector
 *     (Iters.c:723-728);
 *   - the cycle loop stops when |e_k| < tol ||D^-1 b|| (Iters.c:740).
 * Reported: iters = Arnoldi steps + 1 and *cycles = cycles run (the
 * reference's log line, Iters.c:778-786).  status 1 is the reference's
 * MAXIT exit (Iters.c:770-774).  norm_r_m is |e_k| of the last cycle. */
int orc_gmres(int n, int nad, const int *ia, const int *ja, const double *al,
              const double *ad, const double *b, double *x, double tol,
              unsigned ncycles, int newX, int storage, int unsym, int nk,
              orc_stats *st, unsigned *cycles_out)
{
  (void)nad;
  matvec_fn matvec = pick(storage, unsym);
  if (!matvec || n <= 0 || nk < 1) return -1;
  const size_t N = (size_t)n;
  double *m = malloc(N * sizeof(double));
  double *g = calloc(N * (size_t)(nk + 1), sizeof(double));
  double *h = calloc((size_t)nk * (nk + 1), sizeof(double));
  double *e = calloc((size_t)nk + 1, sizeof(double));
  double *c = calloc((size_t)nk, sizeof(double));
  double *s = calloc((size_t)nk, sizeof(double));
  double *y = calloc((size_t)nk, sizeof(double));
  if (!m || !g || !h || !e || !c || !s || !y) {
    free(m); free(g); free(h); free(e); free(c); free(s); free(y);
    return -2;
  }
#define H(i, j) h[(size_t)(i) * nk + (j)]
  for (int i = 0; i < n; i++) m[i] = 1.0 / ad[i];
  const double t0 = now();
  if (newX) for (int i = 0; i < n; i++) x[i] = 0.0;
  for (int i = 0; i < n; i++) g[i] = b[i] * m[i];
  const double norm_b = sqrt(orc_dot(g, g, n));
  const double conv = tol * norm_b;
  unsigned l;
  unsigned nit = 0;
  int ni = 0;
  for (l = 0; l < ncycles; l++) {
    matvec(n, ia, ja, al, ad, x, g);
    for (int i = 0; i < n; i++) g[i] = (b[i] - g[i]) * m[i];
    e[0] = sqrt(orc_dot(g, g, n));
    double tmp = 1.0 / e[0];
    for (int i = 0; i < n; i++) g[i] *= tmp;
    for (ni = 0; ni < nk; ni++, nit++) {
      double *v = g + (size_t)ni * N, *w = g + (size_t)(ni + 1) * N;
      matvec(n, ia, ja, al, ad, v, w);
      for (int i = 0; i < n; i++) w[i] *= m[i];
      for (int j = 0; j <= ni; j++) {
        const double *gj = g + (size_t)j * N;
        const double beta = orc_dot(w, gj, n);
        for (int i = 0; i < n; i++) w[i] -= beta * gj[i];
        H(j, ni) = beta;
      }
      const double nrm = sqrt(orc_dot(w, w, n));
      H(ni + 1, ni) = nrm;
      tmp = 1.0 / nrm;
      for (int i = 0; i < n; i++) w[i] *= tmp;
      for (int j = 0; j < ni; j++) {
        const double h1 = H(j, ni), h2 = H(j + 1, ni);
        H(j, ni) = c[j] * h1 + s[j] * h2;
        H(j + 1, ni) = -s[j] * h1 + c[j] * h2;
      }
      const double h1 = H(ni, ni), h2 = H(ni + 1, ni);
      const double r = sqrt(h1 * h1 + h2 * h2);
      tmp = 1.0 / r;
      c[ni] = h1 * tmp;
      s[ni] = h2 * tmp;
      H(ni, ni) = r;
      H(ni + 1, ni) = 0.0;
      e[ni + 1] = -s[ni] * e[ni];
      e[ni] = c[ni] * e[ni];
    }
    ni--;
    y[ni] = e[ni] / H(ni, ni);
    for (int i = ni - 1; i >= 0; i--) {
      y[i] = 0.0;
      for (int j = i + 1; j < ni; j++) y[i] -= H(i, j) * y[j];
      y[i] = (y[i] + e[i]) / H(i, i);
    }
    for (int j = 0; j < ni; j++) {
      const double yj = y[j];
      const double *gj = g + (size_t)j * N;
      for (int i = 0; i < n; i++) x[i] += gj[i] * yj;
    }
    if (fabs(e[ni + 1]) < conv) break;
  }
  const double last = fabs(e[ni + 1]);
  matvec(n, ia, ja, al, ad, x, g);
  const double xKx = orc_dot(x, g, n);
  const double norm = sqrt(orc_dot(x, x, n));
  for (int i = 0; i < n; i++) g[N + i] = b[i] - g[i];
  const double norm_r = sqrt(orc_dot(g + N, g + N, n));
  const double t1 = now();
#undef H
  if (st) {
    st->status = (l == ncycles) ? 1 : 0;
    st->iters = nit + 1;
    st->xKx = xKx; st->norm_x = norm; st->norm_r = norm_r;
    st->norm_r_m = last; st->norm_b = norm_b; st->conv = conv;
    st->time_s = t1 - t0;
  }
  if (cycles_out) *cycles_out = (l == ncycles) ? l : l + 1;
  free(m); free(g); free(h); free(e); free(c); free(s); free(y);
  return 0;
}

/* BiCGSTAB(2) with right Jacobi preconditioning, reference
 * src/Bicgstab.c:1172-1378 (pbicgstabl2), driven as src/Solv.c:217-295:
 * conv = tol ||b||; per iteration an even and an odd BiCG step (two SpMVs
 * each) and the GCR(2) minimisation over (s, t); the stop test after the
 * double step, on the recurrence residual.  norm_r_m holds that recurrence
 * ||r|| at the end (the true ||b - Ax|| is norm_r), so the two can be
 * compared.  status 1 is the reference's MAXIT exit. */
int orc_pbicgstabl2(int n, int nad, const int *ia, const int *ja,
                    const double *al, const double *ad, const double *b,
                    double *x, double tol, unsigned maxit, int newX,
                    int storage, int unsym, orc_stats *st)
{
  (void)nad;
  matvec_fn matvec = pick(storage, unsym);
  if (!matvec || n <= 0) return -1;
  const size_t N = (size_t)n;
  double *buf = calloc(N * 12, sizeof(double));
  if (!buf) return -2;
  double *m = buf, *t = buf + N, *v = buf + 2 * N, *r = buf + 3 * N;
  double *u = buf + 4 * N, *r0 = buf + 5 * N, *w = buf + 6 * N;
  double *s = buf + 7 * N, *p = buf + 8 * N, *h = buf + 9 * N, *z = buf + 10 * N;
  for (int i = 0; i < n; i++) m[i] = 1.0 / ad[i];
  const double t0 = now();
  if (newX) for (int i = 0; i < n; i++) x[i] = 0.0;
  double d = orc_dot(b, b, n);
  const double norm_b = sqrt(d);
  const double conv = tol * norm_b;
  matvec(n, ia, ja, al, ad, x, t);
  for (int i = 0; i < n; i++) {
    r0[i] = b[i] - t[i];
    r[i] = r0[i];
    u[i] = 0.0;
  }
  double rr0 = 1.0, alpha = 0.0, omega2 = 1.0, omega1, rr1, beta, gamma;
  unsigned j;
  for (j = 0; j < maxit; j++) {
    rr0 = -omega2 * rr0;
    /* even BiCG step */
    rr1 = orc_dot(r, r0, n);
    beta = alpha * rr1 / rr0;
    rr0 = rr1;
    for (int i = 0; i < n; i++) {
      u[i] = r[i] - beta * u[i];
      z[i] = u[i] * m[i];
    }
    matvec(n, ia, ja, al, ad, z, v);
    gamma = orc_dot(v, r0, n);
    alpha = rr0 / gamma;
    for (int i = 0; i < n; i++) {
      r[i] -= alpha * v[i];
      p[i] = r[i] * m[i];
    }
    matvec(n, ia, ja, al, ad, p, s);
    for (int i = 0; i < n; i++) x[i] += alpha * z[i];
    /* odd BiCG step */
    rr1 = orc_dot(s, r0, n);
    beta = alpha * rr1 / rr0;
    rr0 = rr1;
    for (int i = 0; i < n; i++) {
      v[i] = s[i] - beta * v[i];
      p[i] = v[i] * m[i];
    }
    matvec(n, ia, ja, al, ad, p, w);
    gamma = orc_dot(w, r0, n);
    alpha = rr0 / gamma;
    for (int i = 0; i < n; i++) {
      u[i] = r[i] - beta * u[i];
      r[i] -= alpha * v[i];
      s[i] -= alpha * w[i];
      p[i] = s[i] * m[i];
      z[i] = u[i] * m[i];
      h[i] = r[i] * m[i];
    }
    matvec(n, ia, ja, al, ad, p, t);
    /* GCR(2) part */
    omega1 = orc_dot(r, s, n);
    const double mi = orc_dot(s, s, n);
    const double ni = orc_dot(s, t, n);
    double tau = orc_dot(t, t, n);
    omega2 = orc_dot(r, t, n);
    tau = tau - ni * ni / mi;
    omega2 = (omega2 - ni * omega1 / mi) / tau;
    omega1 = (omega1 - ni * omega2) / mi;
    for (int i = 0; i < n; i++) {
      x[i] += omega1 * h[i] + omega2 * p[i] + alpha * z[i];
      r[i] -= (omega1 * s[i] + omega2 * t[i]);
    }
    d = orc_dot(r, r, n);
    if (sqrt(d) < conv) break;
    for (int i = 0; i < n; i++) u[i] -= (omega1 * v[i] + omega2 * w[i]);
  }
  const double rec = sqrt(fabs(d));
  matvec(n, ia, ja, al, ad, x, z);
  const double xKx = orc_dot(x, z, n);
  const double norm = sqrt(orc_dot(x, x, n));
  for (int i = 0; i < n; i++) r[i] = b[i] - z[i];
  const double nr = sqrt(orc_dot(r, r, n));
  const double t1 = now();
  if (st) {
    st->status = j == maxit ? 1 : 0;
    st->iters = j == maxit ? maxit : j + 1;
    st->xKx = xKx; st->norm_x = norm; st->norm_r = nr; st->norm_r_m = rec;
    st->norm_b = norm_b; st->conv = conv; st->time_s = t1 - t0;
  }
  free(buf);
  return 0;
}

/* Plane rotation of src/Iters.c:1171-1208 (symOrtho): c, s, r with
 * [c s; -s c] [a; b] = [r; 0].  Kept as written: the third test compares
 * |b| with a itself, and when none of the four cases holds (a == b > 0) the
 * outputs keep the values they had. */
static void sym_ortho(double a, double b, double *c, double *s, double *r)
{
  const double ma = fabs(a), mb = fabs(b);
  const double sa = a > 0.0 ? 1.0 : -1.0, sb = b > 0.0 ? 1.0 : -1.0;
  if (b == 0.0) {
    *s = 0.0;
    *r = ma;
    *c = a == 0.0 ? 1.0 : sa;
  } else if (a == 0.0) {
    *c = 0.0;
    *s = sb;
    *r = mb;
  } else if (mb > a) {
    const double t = a / b;
    *s = sb / sqrt(1.0 + t * t);
    *c = *s * t;
    *r = b / *s;
  } else if (ma > mb) {
    const double t = b / a;
    *c = sa / sqrt(1.0 + t * t);
    *s = *c * t;
    *r = a / *c;
  }
}

/* Jacobi-preconditioned MINRES, reference src/Iters.c:309-555 (pminres),
 * driven as src/Solv.c:380-445: conv = tol ||D^-1/2 b||; the Lanczos
 * vectors v, z = D^-1 v with beta = sqrt(z.v); the QR step by the rotation
 * above; the stop test on the recurrence |s| ||r||.  Reported iterations:
 * the loop index + 1 at the break; at maxit the reference prints a warning
 * and returns (no exit), status 1 here, iters = maxit. */
int orc_minres(int n, int nad, const int *ia, const int *ja, const double *al,
               const double *ad, const double *b, double *x, double tol,
               unsigned maxit, int newX, int storage, int unsym,
               orc_stats *st)
{
  (void)nad;
  matvec_fn matvec = pick(storage, unsym);
  if (!matvec || n <= 0) return -1;
  const size_t N = (size_t)n;
  double *buf = calloc(N * 10, sizeof(double));
  if (!buf) return -2;
  double *m = buf, *v0 = buf + N, *v = buf + 2 * N, *w = buf + 3 * N;
  double *w0 = buf + 4 * N, *w00 = buf + 5 * N, *z = buf + 6 * N;
  double *z0 = buf + 7 * N, *p = buf + 8 * N;
  for (int i = 0; i < n; i++) m[i] = 1.0 / ad[i];
  const double t0 = now();
  double ro1 = 1.0, ro2, ro3;
  if (newX) for (int i = 0; i < n; i++) x[i] = 0.0;
  const double norm_b = sqrt(orc_dot(b, b, n));
  for (int i = 0; i < n; i++) z[i] = b[i] * sqrt(m[i]);
  const double conv = tol * sqrt(orc_dot(z, z, n));
  matvec(n, ia, ja, al, ad, x, z);
  for (int i = 0; i < n; i++) {
    v[i] = b[i] - z[i];
    v0[i] = w0[i] = w00[i] = 0.0;
    z0[i] = v[i] * m[i];
  }
  double beta = sqrt(fabs(orc_dot(v, z0, n)));
  double norm_r = beta, neta = beta, beta_old = beta;
  double c_old = 1.0, c = 1.0, s_old = 0.0, s = 0.0;
  unsigned j;
  for (j = 0; j < maxit; j++) {
    matvec(n, ia, ja, al, ad, z0, p);
    double tmp1 = 1.0 / (beta * beta);
    const double alpha = orc_dot(z0, p, n) * tmp1;
    tmp1 = 1.0 / beta;
    double tmp2 = -alpha / beta, tmp3 = -beta / beta_old;
    for (int i = 0; i < n; i++) {
      p[i] = tmp1 * p[i] + tmp2 * v[i] + tmp3 * v0[i];
      z[i] = p[i] * m[i];
    }
    beta_old = beta;
    beta = sqrt(orc_dot(z, p, n));
    const double delta = c * alpha - c_old * s * beta_old;
    ro2 = s * alpha + c_old * c * beta_old;
    ro3 = s_old * beta_old;
    c_old = c;
    s_old = s;
    sym_ortho(delta, beta, &c, &s, &ro1);
    tmp1 = 1.0 / (beta_old * ro1);
    tmp2 = -ro2 / ro1;
    tmp3 = -ro3 / ro1;
    const double tmp4 = c * neta;
    for (int i = 0; i < n; i++) {
      w[i] = tmp1 * z0[i] + tmp2 * w0[i] + tmp3 * w00[i];
      x[i] += tmp4 * w[i];
      v0[i] = v[i];
      v[i] = p[i];
      z0[i] = z[i];
      w00[i] = w0[i];
      w0[i] = w[i];
    }
    norm_r = fabs(s) * norm_r;
    if (norm_r < conv) break;
    neta = -s * neta;
  }
  matvec(n, ia, ja, al, ad, x, z);
  const double xKx = orc_dot(x, z, n);
  const double norm = sqrt(orc_dot(x, x, n));
  for (int i = 0; i < n; i++) {
    v[i] = b[i] - z[i];
    z[i] = v[i] * sqrt(fabs(m[i]));
  }
  const double nr = sqrt(orc_dot(v, v, n));
  const double nrm = sqrt(orc_dot(z, z, n));
  const double t1 = now();
  if (st) {
    st->status = j == maxit ? 1 : 0;
    st->iters = j == maxit ? maxit : j + 1;
    st->xKx = xKx; st->norm_x = norm; st->norm_r = nr; st->norm_r_m = nrm;
    st->norm_b = norm_b; st->conv = conv; st->time_s = t1 - t0;
  }
  free(buf);
  return 0;
}

/* ---- all-cores analogue (OpenMP over full rows) ---------------------------
 * Matrix given as full rows: rp[n+1], cj[], cv[] off-diagonal, ad[] diagonal
 * (build it with orc_full_rows).  Same recurrences as orc_pcg. */

int orc_full_rows(int storage, int unsym, int n, const int *ia, const int *ja,
                  const double *al, int64_t *rp, int *cj, double *cv)
{
  /* counts */
  memset(rp, 0, sizeof(int64_t) * ((size_t)n + 1));
  const int csrc = storage == ORC_CSRC || storage == ORC_CSRCCOO;
  const int nad = storage == ORC_ELLPACK ? 0 : ia[n];
  for (int i = 0; i < n; i++) {
    if (storage == ORC_ELLPACK) { rp[i + 1] += ia[0]; continue; }
    for (int k = ia[i]; k < ia[i + 1]; k++) {
      const int j = ja[k];
      if (storage == ORC_CSR && j == i) continue;
      rp[i + 1]++;
      if (!unsym || csrc) rp[j + 1]++;
    }
  }
  for (int i = 0; i < n; i++) rp[i + 1] += rp[i];
  int64_t *pos = malloc(sizeof(int64_t) * (size_t)n);
  if (!pos) return -1;
  for (int i = 0; i < n; i++) pos[i] = rp[i];
  for (int i = 0; i < n; i++) {
    if (storage == ORC_ELLPACK) {
      const int s = ia[0];
      for (int f = 0; f < s; f++) {
        cj[pos[i]] = ja[i * s + f]; cv[pos[i]++] = al[i * s + f];
      }
      continue;
    }
    for (int k = ia[i]; k < ia[i + 1]; k++) {
      const int j = ja[k];
      if (storage == ORC_CSR && j == i) continue;
      cj[pos[i]] = j; cv[pos[i]++] = al[k];
      if (csrc && unsym) { cj[pos[j]] = i; cv[pos[j]++] = al[nad + k]; }
      else if (!unsym) { cj[pos[j]] = i; cv[pos[j]++] = al[k]; }
    }
  }
  free(pos);
  return 0;
}

static void mv_full_omp(int n, const int64_t *rp, const int *cj,
                        const double *cv, const double *ad, const double *x,
                        double *y)
{
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    double t = ad[i] * x[i];
    for (int64_t k = rp[i]; k < rp[i + 1]; k++) t += cv[k] * x[cj[k]];
    y[i] = t;
  }
}

static double dot_omp(const double *a, const double *b, int n)
{
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

int orc_pcg_omp(int n, const int64_t *rp, const int *cj, const double *cv,
                const double *ad, const double *b, double *x, double tol,
                unsigned maxit, int nthreads, orc_stats *st)
{
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  double *m = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n);
  double *r = malloc(sizeof(double) * n), *p = malloc(sizeof(double) * n);
  if (!m || !z || !r || !p) { free(m); free(z); free(r); free(p); return -2; }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) { m[i] = 1.0 / ad[i]; x[i] = 0.0; z[i] = 0.0; r[i] = 0.0; p[i] = 0.0; }
  const double t0 = now();
  unsigned j;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) z[i] = b[i] * m[i];
  double d = dot_omp(b, z, n);
  const double norm_b = sqrt(fabs(d)), conv = tol * norm_b;
  mv_full_omp(n, rp, cj, cv, ad, x, z);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) { r[i] = b[i] - z[i]; z[i] = r[i] * m[i]; p[i] = z[i]; }
  d = dot_omp(r, z, n);
  for (j = 0; j < maxit; j++) {
    mv_full_omp(n, rp, cj, cv, ad, p, z);
    const double alpha = d / dot_omp(z, p, n);
    double di = 0.0;
#pragma omp parallel for reduction(+ : di) schedule(static)
    for (int i = 0; i < n; i++) {
      x[i] += alpha * p[i];
      r[i] -= alpha * z[i];
      z[i] = r[i] * m[i];
      di += r[i] * z[i];
    }
    const double beta = di / d;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
    d = di;
    if (sqrt(fabs(d)) < conv) break;
  }
  const double t1 = now();
  mv_full_omp(n, rp, cj, cv, ad, x, z);
  double rr = 0.0;
  for (int i = 0; i < n; i++) { const double ri = b[i] - z[i]; rr += ri * ri; }
  if (st) {
    memset(st, 0, sizeof *st);
    st->status = (j == maxit) ? 1 : 0;
    st->iters = j + 1;
    st->norm_r = sqrt(rr);
    st->norm_b = norm_b; st->conv = conv;
    st->time_s = t1 - t0;
  }
  free(m); free(z); free(r); free(p);
  return 0;
}

int orc_max_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
