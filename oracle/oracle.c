/*
 * CPU oracle for the pressure-Poisson symmetric-solver path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it.
 * The product (libaphpois) never links, calls or falls back to it.
 *
 * It is a plain-C restatement of the reference's algorithms, written from the
 * reference's behaviour (cselab/aphros, src/linear and src/solver):
 *
 *   orc_assemble       pressure system of Proj::GetFlux/GetFluxSum with the
 *                      harmonic face density (src/solver/proj.ipp:343-383,
 *                      src/solver/approx_eb.h:351-363,
 *                      src/solver/approx_eb.ipp:835-837,1435-1446,
 *                      src/geom/mesh.h:575-579)
 *   orc_apply          Expr evaluation e0*x + sum_q e[1+q]*x[nb(q)] (+ e7)
 *                      (src/linear/linear.ipp:49-56,65-72,
 *                      src/solver/approx_eb.h:328-336)
 *   orc_conjugate      SolverConjugate::Imp::Solve (src/linear/linear.ipp:24-129)
 *   orc_jacobi         SolverJacobi::Imp::Solve (src/linear/linear.ipp:158-217)
 *   orc_pcg_jacobi     textbook Jacobi-preconditioned CG in FP64 with a true
 *                      relative residual stop; the reference has no
 *                      preconditioned CPU CG, this is the FP64 yardstick for
 *                      the preconditioned GPU forms.
 *   orc_true_relres    ||b - A x||_2 / ||b||_2 with b = -e7, in FP64.
 *
 * Grid: nx*ny*nz cells, x fastest, fully periodic (the reference test and
 * default config are periodic: src/distr/distrbasic.cpp:66-83).  Neighbour
 * q = 2*d + side, side 0 = minus (src/geom/mesh.ipp:100-107).  Expr AoS with 8
 * doubles per cell: e0 (diagonal), e1..e6 (xm, xp, ym, yp, zm, zp), e7
 * (constant) (src/geom/mesh.h:482-485).
 *
 * Reductions are summed in cell order over the whole grid, which is what the
 * reference does for one block (src/distr/local.ipp:153-172 with a single
 * block).  Per-cell arithmetic follows the reference's operation order so that
 * a one-block reference run reproduces bit for bit.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
  int nx, ny, nz;
} grid_t;

static inline size_t idx(const grid_t* g, int x, int y, int z) {
  return (size_t)x + (size_t)g->nx * ((size_t)y + (size_t)g->ny * (size_t)z);
}

static inline int wrap(int i, int n) {
  return i < 0 ? i + n : (i >= n ? i - n : i);
}

/* neighbour indices in reference order: xm, xp, ym, yp, zm, zp */
static inline void neighbours(const grid_t* g, int x, int y, int z, size_t nb[6]) {
  nb[0] = idx(g, wrap(x - 1, g->nx), y, z);
  nb[1] = idx(g, wrap(x + 1, g->nx), y, z);
  nb[2] = idx(g, x, wrap(y - 1, g->ny), z);
  nb[3] = idx(g, x, wrap(y + 1, g->ny), z);
  nb[4] = idx(g, x, y, wrap(z - 1, g->nz));
  nb[5] = idx(g, x, y, wrap(z + 1, g->nz));
}

/*
 * Pressure system from a cell density field (dt = 1, zero face fluxes, volume
 * source f).  Follows the reference's floating-point sequence:
 *   h = 1/n (cell_size = extent/n), vol = h*h*h, area = vol/h
 *   (src/geom/mesh.ipp:84-87);
 *   inverse density per cell, face value (inv[cp] + inv[cm]) * 0.5, then
 *   dens_f = 1/that (InterpolateHarmonic, approx_eb.h:351-363, Interpolate
 *   approx_eb.ipp:835-837);
 *   GradientImplicit ExprFace(-1/h, 1/h, 0) (approx_eb.ipp:1435-1446);
 *   scaled by s = -area/dens_f*dt (proj.ipp:356-358);
 *   AppendExpr with the outward factor (proj.ipp:367-383, mesh.h:575-579);
 *   constant term -= f*vol (proj.ipp:379).
 * The cell-loop order of AppendExpr is q = 0..5.
 * sign > 0 gives the Proj convention (e0 > 0); sign < 0 negates every
 * coefficient (the SolvePoisson / t.linear convention, e0 < 0) and the
 * constant term becomes +f*vol.
 */
void orc_assemble(int nx, int ny, int nz, double extent, const double* rho,
                  const double* f, int sign, double* expr) {
  grid_t g = {nx, ny, nz};
  const double h = extent / (double)nx;
  const double vol = h * h * h;
  const double area = vol / h;
  const double a = 1.0 / h;
  const size_t n = (size_t)nx * ny * nz;
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const size_t c = idx(&g, x, y, z);
        size_t nb[6];
        neighbours(&g, x, y, z, nb);
        double e[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int q = 0; q < 6; ++q) {
          const int plus = q & 1;
          /* face cells: cm is the lower-index side, cp the upper */
          const size_t cm = plus ? c : nb[q];
          const size_t cp = plus ? nb[q] : c;
          const double inv = (1.0 / rho[cp] + 1.0 / rho[cm]) * 0.5;
          const double dens = 1.0 / inv;
          const double s = -area / dens * 1.0;
          double v0 = -a * s, v1 = a * s, v2 = 0.0;
          if (!plus) { /* outward factor -1 */
            v0 = -v0;
            v1 = -v1;
            v2 = -v2;
          }
          e[0] += plus ? v0 : v1;
          e[1 + q] += plus ? v1 : v0;
          e[7] += v2;
        }
        e[7] -= f[c] * vol;
        for (int k = 0; k < 8; ++k) {
          expr[c * 8 + k] = sign > 0 ? e[k] : -e[k];
        }
      }
  (void)n;
}

/* y = A x (+ e7 if with_const), reference order: ((e0*x_c [+ e7]) + e1*x_xm) + ... */
void orc_apply(const double* expr, int nx, int ny, int nz, const double* x,
               double* y, int with_const) {
  grid_t g = {nx, ny, nz};
  for (int zz = 0; zz < nz; ++zz)
    for (int yy = 0; yy < ny; ++yy)
      for (int xx = 0; xx < nx; ++xx) {
        const size_t c = idx(&g, xx, yy, zz);
        const double* e = expr + c * 8;
        size_t nb[6];
        neighbours(&g, xx, yy, zz, nb);
        double u = x[c] * e[0];
        if (with_const) u += e[7];
        for (int q = 0; q < 6; ++q) u += x[nb[q]] * e[1 + q];
        y[c] = u;
      }
}

/*
 * SolverConjugate (src/linear/linear.ipp:24-129).
 * h3 = GetCellSize().prod().  x0 may be NULL (zero guess) or equal x.
 * hist (optional, length hist_len) receives info.residual after each iteration.
 * Returns info.iter; *res_out = info.residual.
 */
int orc_conjugate(const double* expr, int nx, int ny, int nz, double h3,
                  const double* x0, double* x, double tol, int miniter,
                  int maxiter, int maxnorm, double* hist, int hist_len,
                  double* res_out) {
  grid_t g = {nx, ny, nz};
  const size_t n = (size_t)nx * ny * nz;
  double* u = (double*)malloc(n * sizeof(double));
  double* r = (double*)malloc(n * sizeof(double));
  double* p = (double*)malloc(n * sizeof(double));
  double* lp = (double*)malloc(n * sizeof(double));
  if (!u || !r || !p || !lp) {
    free(u); free(r); free(p); free(lp);
    return -1;
  }
  if (x0) memcpy(u, x0, n * sizeof(double));
  else memset(u, 0, n * sizeof(double));
  /* init: r = -(A u + e7)  (linear.ipp:49-56) */
  orc_apply(expr, nx, ny, nz, u, r, 1);
  for (size_t c = 0; c < n; ++c) r[c] = -r[c];
  memcpy(p, r, n * sizeof(double));
  int iter = 0;
  double residual = 0;
  for (;;) {
    /* iter (linear.ipp:64-82) */
    orc_apply(expr, nx, ny, nz, p, lp, 0);
    double dot_r_prev = 0, dot_p_lp = 0;
    for (size_t c = 0; c < n; ++c) {
      dot_r_prev += r[c] * r[c];
      dot_p_lp += p[c] * lp[c];
    }
    /* iter2 (linear.ipp:83-95) */
    const double alpha = dot_r_prev / (dot_p_lp + 1e-100);
    double dot_r = 0, max_r = 0;
    for (size_t c = 0; c < n; ++c) {
      u[c] += alpha * p[c];
      r[c] -= alpha * lp[c];
      dot_r += r[c] * r[c];
      const double ar = fabs(r[c]);
      max_r = max_r > ar ? max_r : ar;
    }
    /* iter3 (linear.ipp:96-101) */
    for (size_t c = 0; c < n; ++c) {
      p[c] = r[c] + (dot_r / (dot_r_prev + 1e-100)) * p[c];
    }
    /* check (linear.ipp:102-114) */
    residual = maxnorm ? max_r / h3 : sqrt(dot_r / h3);
    ++iter;
    if (hist && iter <= hist_len) hist[iter - 1] = residual;
    if (iter >= miniter && (iter > maxiter || residual < tol)) break;
  }
  memcpy(x, u, n * sizeof(double));
  if (res_out) *res_out = residual;
  free(u); free(r); free(p); free(lp);
  (void)g;
  return iter;
}

/*
 * SolverJacobi (src/linear/linear.ipp:158-217).
 * x_new = -(e7 + sum_q e[1+q] x[nb]) / e0; residual = max |x_new - x|.
 */
int orc_jacobi(const double* expr, int nx, int ny, int nz, const double* x0,
               double* x, double tol, int miniter, int maxiter, double* hist,
               int hist_len, double* res_out) {
  grid_t g = {nx, ny, nz};
  const size_t n = (size_t)nx * ny * nz;
  double* u = (double*)malloc(n * sizeof(double));
  double* un = (double*)malloc(n * sizeof(double));
  if (!u || !un) {
    free(u); free(un);
    return -1;
  }
  if (x0) memcpy(u, x0, n * sizeof(double));
  else memset(u, 0, n * sizeof(double));
  int iter = 0;
  double residual = 0;
  for (;;) {
    double maxdiff = 0;
    for (int zz = 0; zz < nz; ++zz)
      for (int yy = 0; yy < ny; ++yy)
        for (int xx = 0; xx < nx; ++xx) {
          const size_t c = idx(&g, xx, yy, zz);
          const double* e = expr + c * 8;
          size_t nb[6];
          neighbours(&g, xx, yy, zz, nb);
          double nondiag = e[7];
          for (int q = 0; q < 6; ++q) nondiag += u[nb[q]] * e[1 + q];
          un[c] = -nondiag / e[0];
          const double d = fabs(un[c] - u[c]);
          maxdiff = maxdiff > d ? maxdiff : d;
        }
    double* t = u; u = un; un = t;
    residual = maxdiff;
    ++iter;
    if (hist && iter <= hist_len) hist[iter - 1] = residual;
    if (iter >= miniter && (iter > maxiter || residual < tol)) break;
  }
  memcpy(x, u, n * sizeof(double));
  if (res_out) *res_out = residual;
  free(u); free(un);
  return iter;
}

/* ||b - A x||_2 / ||b||_2 with b = -e7 (FP64, cell-order sums). */
double orc_true_relres(const double* expr, int nx, int ny, int nz,
                       const double* x) {
  const size_t n = (size_t)nx * ny * nz;
  double* y = (double*)malloc(n * sizeof(double));
  if (!y) return -1;
  orc_apply(expr, nx, ny, nz, x, y, 1); /* A x + e7 = A x - b */
  double rr = 0, bb = 0;
  for (size_t c = 0; c < n; ++c) {
    rr += y[c] * y[c];
    bb += expr[c * 8 + 7] * expr[c * 8 + 7];
  }
  free(y);
  return bb > 0 ? sqrt(rr / bb) : sqrt(rr);
}

/*
 * Jacobi-preconditioned CG in FP64 (textbook PCG, M = diag(A)), zero or given
 * initial guess, stopped when the recursive residual satisfies
 * ||r||_2 <= tol * ||b||_2 and then confirmed on the true residual (the loop
 * continues with a tighter target if the true residual is above tol).
 * OpenMP-parallel when built with -fopenmp (sums then differ from cell order
 * in the last bits; iteration counts are what the oracle is used for here).
 * If remove_mean, the mean of x is removed at the end (periodic nullspace).
 * Returns the iteration count; *relres_out = true relative residual.
 * If maxiter < 0 the loop runs exactly -maxiter iterations (timing sample).
 */
int orc_pcg_jacobi(const double* expr, int nx, int ny, int nz, const double* x0,
                   double* x, double tol, int maxiter, int remove_mean,
                   double* relres_out) {
  grid_t g = {nx, ny, nz};
  const long long n = (long long)nx * ny * nz;
  double* r = (double*)malloc(n * sizeof(double));
  double* z = (double*)malloc(n * sizeof(double));
  double* p = (double*)malloc(n * sizeof(double));
  double* q = (double*)malloc(n * sizeof(double));
  if (!r || !z || !p || !q) {
    free(r); free(z); free(p); free(q);
    return -1;
  }
  const int fixed = maxiter < 0;
  const int itmax = fixed ? -maxiter : maxiter;
  if (x0 && x0 != x) memcpy(x, x0, n * sizeof(double));
  if (!x0) memset(x, 0, n * sizeof(double));
  double bb = 0;
#pragma omp parallel for reduction(+ : bb) schedule(static)
  for (long long c = 0; c < n; ++c) bb += expr[c * 8 + 7] * expr[c * 8 + 7];
  const double bnorm = sqrt(bb);
  double target = tol;
  int iter = 0;
  double relres = 1;
  for (int restart = 0; restart < 8; ++restart) {
    /* r = b - A x = -(A x + e7) */
#pragma omp parallel for schedule(static)
    for (int zz = 0; zz < nz; ++zz)
      for (int yy = 0; yy < ny; ++yy)
        for (int xx = 0; xx < nx; ++xx) {
          const size_t c = idx(&g, xx, yy, zz);
          const double* e = expr + c * 8;
          size_t nb[6];
          neighbours(&g, xx, yy, zz, nb);
          double u = x[c] * e[0] + e[7];
          for (int k = 0; k < 6; ++k) u += x[nb[k]] * e[1 + k];
          r[c] = -u;
        }
    double rz = 0, rr = 0;
#pragma omp parallel for reduction(+ : rz, rr) schedule(static)
    for (long long c = 0; c < n; ++c) {
      z[c] = r[c] / expr[c * 8];
      p[c] = z[c];
      rz += r[c] * z[c];
      rr += r[c] * r[c];
    }
    relres = bnorm > 0 ? sqrt(rr) / bnorm : sqrt(rr);
    if (!fixed && relres <= tol) break;
    while (fixed ? iter < itmax : (iter < itmax && sqrt(rr) > target * bnorm)) {
      double pq = 0;
#pragma omp parallel for reduction(+ : pq) schedule(static)
      for (int zz = 0; zz < nz; ++zz)
        for (int yy = 0; yy < ny; ++yy)
          for (int xx = 0; xx < nx; ++xx) {
            const size_t c = idx(&g, xx, yy, zz);
            const double* e = expr + c * 8;
            size_t nb[6];
            neighbours(&g, xx, yy, zz, nb);
            double u = p[c] * e[0];
            for (int k = 0; k < 6; ++k) u += p[nb[k]] * e[1 + k];
            q[c] = u;
            pq += p[c] * u;
          }
      const double alpha = rz / pq;
      double rz_new = 0;
      rr = 0;
#pragma omp parallel for reduction(+ : rz_new, rr) schedule(static)
      for (long long c = 0; c < n; ++c) {
        x[c] += alpha * p[c];
        r[c] -= alpha * q[c];
        z[c] = r[c] / expr[c * 8];
        rz_new += r[c] * z[c];
        rr += r[c] * r[c];
      }
      const double beta = rz_new / rz;
      rz = rz_new;
#pragma omp parallel for schedule(static)
      for (long long c = 0; c < n; ++c) p[c] = z[c] + beta * p[c];
      ++iter;
    }
    if (fixed) break;
    relres = orc_true_relres(expr, nx, ny, nz, x);
    if (relres <= tol || iter >= itmax) break;
    target *= 0.5 * tol / relres;
  }
  if (remove_mean) {
    double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (long long c = 0; c < n; ++c) s += x[c];
    const double mean = s / (double)n;
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < n; ++c) x[c] -= mean;
  }
  if (relres_out) *relres_out = fixed ? relres : orc_true_relres(expr, nx, ny, nz, x);
  free(r); free(z); free(p); free(q);
  return iter;
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
