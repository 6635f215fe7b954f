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
 *   orc_assemble_flux  the same with face volume fluxes, time step and
 *                      source (GetFlux / GetFluxSum, src/solver/proj.ipp:343-383,
 *                      AppendExpr src/geom/mesh.h:575-579)
 *   orc_flux_correct   corrected face fluxes after the solve (proj.ipp:399-403,
 *                      UEB::Eval src/solver/approx_eb.h:338-344)
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
 *   orc_mg_apply,      the multigrid V-cycle preconditioner and CG with it in
 *   orc_mgpcg          FP64, the yardstick for the GPU's multigrid PCG (the
 *                      reference's symmetric plugin has no preconditioner).
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

/*
 * The projection's pressure system with face volume fluxes and a time step
 * (Proj::GetFlux / GetFluxSum, src/solver/proj.ipp:343-383).  Per face f
 * between cm (lower) and cp (upper):
 *   ffe = ExprFace(-1/h, 1/h, 0) * s,  s = -area / dens_f * dt
 *   ffe[2] += ffv[f]                                  (proj.ipp:356-358)
 * and per cell, for q = 0..5, v = ffe * outward(c, q) appended with
 *   sum[0] += v[1 - q%2]; sum[1+q] += v[q%2]; sum[7] += v[2]
 * (src/geom/mesh.h:575-579), then sum[7] -= src * vol (proj.ipp:379).
 * ffv holds three face fields of nx*ny*nz values (x, y, z faces); entry c of
 * direction d is the lower face of cell c (between c - e_d and c).
 * src may be NULL (no source).  sign < 0 negates every coefficient.
 * With fk (and rho NULL) the face denominator is given per face (same layout
 * as ffv) instead of the harmonic density: the pressure correction of SIMPLE
 * (src/solver/simple.ipp:214-226, ffe *= -area / ffk; pass dt = 1).
 */
static double face_den(const double* rho, const double* fk, size_t n, int d, size_t cm,
                       size_t cp) {
  if (fk) return fk[(size_t)d * n + cp];
  const double inv = (1.0 / rho[cp] + 1.0 / rho[cm]) * 0.5;
  return 1.0 / inv;
}

void orc_assemble_flux(int nx, int ny, int nz, double cell, double dt, const double* rho,
                       const double* fk, const double* ffv, const double* src, int sign,
                       double* expr) {
  grid_t g = {nx, ny, nz};
  const double h = cell;
  const double vol = h * h * h;
  const double area = vol / h;
  const double a = 1.0 / h;
  const size_t n = (size_t)nx * ny * nz;
#pragma omp parallel for schedule(static)
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const size_t c = idx(&g, x, y, z);
        size_t nb[6];
        neighbours(&g, x, y, z, nb);
        double e[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int q = 0; q < 6; ++q) {
          const int plus = q & 1;
          const int d = q >> 1;
          const size_t cm = plus ? c : nb[q];
          const size_t cp = plus ? nb[q] : c;
          const double dens = face_den(rho, fk, n, d, cm, cp);
          const double s = -area / dens * dt;
          double v0 = -a * s, v1 = a * s, v2 = 0.0 * s;
          v2 += ffv[(size_t)d * n + cp];  /* the face's flux: lower face of cp */
          if (!plus) {
            v0 = -v0;
            v1 = -v1;
            v2 = -v2;
          }
          e[0] += plus ? v0 : v1;
          e[1 + q] += plus ? v1 : v0;
          e[7] += v2;
        }
        if (src) e[7] -= src[c] * vol;
        for (int k = 0; k < 8; ++k) expr[c * 8 + k] = sign > 0 ? e[k] : -e[k];
      }
}

/*
 * Corrected face fluxes after the pressure solve (Proj::Project "fluxes",
 * src/solver/proj.ipp:399-403; UEB::Eval of an ExprFace,
 * src/solver/approx_eb.h:338-344):  out[f] = p[cm] * ffe[0] + p[cp] * ffe[1]
 * + ffe[2] with ffe of orc_assemble_flux.  Same face-field layout.
 */
void orc_flux_correct(int nx, int ny, int nz, double cell, double dt, const double* rho,
                      const double* fk, const double* ffv, const double* p, double* out) {
  grid_t g = {nx, ny, nz};
  const double h = cell;
  const double vol = h * h * h;
  const double area = vol / h;
  const double a = 1.0 / h;
  const size_t n = (size_t)nx * ny * nz;
#pragma omp parallel for schedule(static)
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const size_t cp = idx(&g, x, y, z);
        size_t nb[6];
        neighbours(&g, x, y, z, nb);
        for (int d = 0; d < 3; ++d) {
          const size_t cm = nb[2 * d];
          const double dens = face_den(rho, fk, n, d, cm, cp);
          const double s = -area / dens * dt;
          const double e0 = -a * s, e1 = a * s;
          double e2 = 0.0 * s;
          e2 += ffv[(size_t)d * n + cp];
          const double t0 = p[cm] * e0;
          const double t1 = p[cp] * e1;
          out[(size_t)d * n + cp] = t0 + t1 + e2;
        }
      }
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

/*
 * Multigrid-preconditioned CG in FP64: the yardstick for the GPU's multigrid
 * PCG (the reference has no such CPU solver).  The preconditioner is one
 * symmetric V-cycle, defined as follows.
 *
 * Operator in conductance form.  With s = sign(e0), the system s*A x = s*b
 * has (s*A u)_c = sum_f k_f (u_c - u_nb) + m_c u_c, k_f = -s * e_plus(c) on
 * the face to the plus neighbour (symmetry gives the minus faces) and
 * m_c = s * (row sum) (0 when |row sum| <= 1e-12 |e0|).
 *
 * Levels.  Level 0 is the grid; level l+1 halves every edge of level l while
 * all edges are even and the coarse grid keeps at least 8 cells (at most 12
 * levels; at least 2, the coarsest at most 4096 cells).  A coarse face
 * conductance is 0.5 * (sum of the 4 fine faces it covers), the coarse mass
 * 0.5 * (sum over the 8 children); D = diagonal of the level operator.
 *
 * V-cycle z = M r on level l (a, b = kMgOmegaA, kMgOmegaB; w = 0.9):
 *   l < top: u = a D^-1 r; z = u + b D^-1 (r - L u);
 *            r_{l+1} = sum over the 2x2x2 children of (r - L z);
 *            z_{l+1} = M r_{l+1};
 *            u = z + p P z_{l+1} (P = injection to the children, p =
 *            kMgProlong, an over-correction of the constant interpolation);
 *            z = u + b D^-1 (r - L u); z = z + a D^-1 (r - L z)
 *   top:     u = w D^-1 r, then `sweeps - 1` more damped Jacobi sweeps
 * Restriction is the transpose of the injection, and pre- and post-smoothing
 * apply the same polynomial of D^-1 L (the error factor (1 - a D^-1 L)
 * (1 - b D^-1 L), a Chebyshev polynomial of degree 2), so M is symmetric.
 * W levels: on the top w levels (l < w, and only where level l+1 is not the
 * coarsest: w <= nlev - 2) the coarse correction is applied twice: with
 * zs = z_{l+1}, r_{l+1} -= L_{l+1} zs, z_{l+1} = M r_{l+1}, and the
 * prolongation adds p (z_{l+1} + zs).  Two symmetric corrections in sequence
 * keep M symmetric.
 *
 * CG: textbook PCG with M, stopped on the recursive residual
 * ||r|| <= target ||b|| and confirmed on the true residual (restart with a
 * tighter target if needed), like orc_pcg_jacobi.  maxiter < 0 runs exactly
 * -maxiter iterations (timing sample).  *setup_s receives the seconds of the
 * level setup.  Returns the iteration count or -1 (allocation), -2 (grid
 * cannot be coarsened).
 */
typedef struct {
  int nx, ny, nz;
  size_t n;
  double *kx, *ky, *kz, *m, *dinv, *r, *z, *t;
} mglev_t;

/* coarsest level: damped Jacobi, weight kMgOmega; other levels: the two
 * weights of the degree-2 Chebyshev polynomial on [0.5, 1.8] of the spectrum
 * of D^-1 L (1 / (1.15 +- 0.65 cos(pi/4))) */
static const double kMgOmega = 0.9;
static const double kMgOmegaA = 0.6212648748965077;
static const double kMgOmegaB = 1.4484764074432;
/* scale of the prolonged coarse correction */
static const double kMgProlong = 1.1;

static double wall_seconds(void) {
#ifdef _OPENMP
  return omp_get_wtime();
#else
  return (double)clock() / CLOCKS_PER_SEC;
#endif
}

/* (L u)_c of level L at (x, y, z) */
static inline double mg_lap(const mglev_t* L, const double* u, int x, int y, int z) {
  const grid_t g = {L->nx, L->ny, L->nz};
  const size_t c = idx(&g, x, y, z);
  const size_t xm = idx(&g, wrap(x - 1, L->nx), y, z), xp = idx(&g, wrap(x + 1, L->nx), y, z);
  const size_t ym = idx(&g, x, wrap(y - 1, L->ny), z), yp = idx(&g, x, wrap(y + 1, L->ny), z);
  const size_t zm = idx(&g, x, y, wrap(z - 1, L->nz)), zp = idx(&g, x, y, wrap(z + 1, L->nz));
  const double uc = u[c];
  double s = L->kx[c] * (uc - u[xp]) + L->kx[xm] * (uc - u[xm]) + L->ky[c] * (uc - u[yp]) +
             L->ky[ym] * (uc - u[ym]) + L->kz[c] * (uc - u[zp]) + L->kz[zm] * (uc - u[zm]);
  if (L->m) s += L->m[c] * uc;
  return s;
}

/* out = u + w D^-1 (r - L u) */
static void mg_sweep(const mglev_t* L, const double* u, double* out, double w) {
#pragma omp parallel for schedule(static)
  for (int z = 0; z < L->nz; ++z)
    for (int y = 0; y < L->ny; ++y)
      for (int x = 0; x < L->nx; ++x) {
        const size_t c = (size_t)x + (size_t)L->nx * ((size_t)y + (size_t)L->ny * z);
        out[c] = u[c] + w * L->dinv[c] * (L->r[c] - mg_lap(L, u, x, y, z));
      }
}

static void mg_vcycle(mglev_t* lev, int l, int top, int sweeps, int w) {
  mglev_t* L = &lev[l];
  const long long n = (long long)L->n;
#pragma omp parallel for schedule(static)
  for (long long c = 0; c < n; ++c)
    L->t[c] = (l == top ? kMgOmega : kMgOmegaA) * L->dinv[c] * L->r[c];
  if (l == top) {
    for (int k = 1; k < sweeps; ++k) {
      mg_sweep(L, L->t, L->z, kMgOmega);
      double* tmp = L->t;
      L->t = L->z;
      L->z = tmp;
    }
    double* tmp = L->t;
    L->t = L->z;
    L->z = tmp;
    return;
  }
  mg_sweep(L, L->t, L->z, kMgOmegaB);
  mglev_t* C = &lev[l + 1];
#pragma omp parallel for schedule(static)
  for (int z = 0; z < C->nz; ++z)
    for (int y = 0; y < C->ny; ++y)
      for (int x = 0; x < C->nx; ++x) {
        double sum = 0;
        for (int d = 0; d < 8; ++d) {
          const int fx = 2 * x + (d & 1), fy = 2 * y + ((d >> 1) & 1), fz = 2 * z + (d >> 2);
          const size_t f = (size_t)fx + (size_t)L->nx * ((size_t)fy + (size_t)L->ny * fz);
          sum += L->r[f] - mg_lap(L, L->z, fx, fy, fz);
        }
        C->r[(size_t)x + (size_t)C->nx * ((size_t)y + (size_t)C->ny * z)] = sum;
      }
  mg_vcycle(lev, l + 1, top, sweeps, w);
  const int second = l < w && l + 1 < top;
  double* zs = NULL;
  if (second) {
    const size_t nc = C->n;
    zs = (double*)malloc(nc * sizeof(double));
    memcpy(zs, C->z, nc * sizeof(double));
    double* rn = (double*)malloc(nc * sizeof(double));
#pragma omp parallel for schedule(static)
    for (int z = 0; z < C->nz; ++z)
      for (int y = 0; y < C->ny; ++y)
        for (int x = 0; x < C->nx; ++x) {
          const size_t c = (size_t)x + (size_t)C->nx * ((size_t)y + (size_t)C->ny * z);
          rn[c] = C->r[c] - mg_lap(C, zs, x, y, z);
        }
    memcpy(C->r, rn, nc * sizeof(double));
    free(rn);
    mg_vcycle(lev, l + 1, top, sweeps, w);
  }
#pragma omp parallel for schedule(static)
  for (int z = 0; z < L->nz; ++z)
    for (int y = 0; y < L->ny; ++y)
      for (int x = 0; x < L->nx; ++x) {
        const size_t c = (size_t)x + (size_t)L->nx * ((size_t)y + (size_t)L->ny * z);
        const size_t cc =
            (size_t)(x >> 1) + (size_t)C->nx * ((size_t)(y >> 1) + (size_t)C->ny * (z >> 1));
        L->t[c] = L->z[c] + kMgProlong * (zs ? C->z[cc] + zs[cc] : C->z[cc]);
      }
  free(zs);
  mg_sweep(L, L->t, L->z, kMgOmegaB);
  mg_sweep(L, L->z, L->t, kMgOmegaA);
  double* tmp = L->t;
  L->t = L->z;
  L->z = tmp;
}

static void mg_free(mglev_t* lev, int nlev) {
  for (int l = 0; l < nlev; ++l) {
    free(lev[l].kx); free(lev[l].ky); free(lev[l].kz); free(lev[l].m);
    free(lev[l].dinv); free(lev[l].r); free(lev[l].z); free(lev[l].t);
  }
}

/* v rounded to float, then to bfloat16 (round to nearest even) */
static double bf16_round(double v) {
  float f = (float)v;
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  memcpy(&f, &u, 4);
  return (double)f;
}

/*
 * Levels of the grid (conductances, masses, D^-1); returns nlev or -1/-2.
 * coef16: the coefficient storage of the GPU's FP32 engines: conductances and
 * D^-1 of every level rounded to bfloat16 (level 0 through the FP32
 * conductances), masses in FP32, coarse sums and D in FP32 arithmetic.
 */
static int mg_setup(const double* expr, int nx, int ny, int nz, int coef16, mglev_t* lev) {
  int nlev = 0;
  int lx = nx, ly = ny, lz = nz;
  for (;;) {
    mglev_t* L = &lev[nlev++];
    memset(L, 0, sizeof(*L));
    L->nx = lx; L->ny = ly; L->nz = lz;
    L->n = (size_t)lx * ly * lz;
    if (nlev >= 12 || ((lx | ly | lz) & 1) || (long long)(lx / 2) * (ly / 2) * (lz / 2) < 8)
      break;
    lx /= 2; ly /= 2; lz /= 2;
  }
  if (nlev < 2 || lev[nlev - 1].n > 4096) return -2;
  const size_t n0 = lev[0].n;
  double sgn = 1;
  for (size_t c = 0; c < n0; ++c)
    if (expr[c * 8] != 0) {
      sgn = expr[c * 8] < 0 ? -1 : 1;
      break;
    }
  int has_m = 0;
  for (size_t c = 0; c < n0 && !has_m; ++c) {
    const double* e = expr + c * 8;
    const double rs = e[0] + e[1] + e[2] + e[3] + e[4] + e[5] + e[6];
    if (fabs(rs) > 1e-12 * fabs(e[0])) has_m = 1;
  }
  for (int l = 0; l < nlev; ++l) {
    mglev_t* L = &lev[l];
    const size_t b = L->n * sizeof(double);
    L->kx = (double*)malloc(b); L->ky = (double*)malloc(b); L->kz = (double*)malloc(b);
    L->m = has_m ? (double*)malloc(b) : NULL;
    L->dinv = (double*)malloc(b); L->r = (double*)malloc(b);
    L->z = (double*)malloc(b); L->t = (double*)malloc(b);
    if (!L->kx || !L->ky || !L->kz || (has_m && !L->m) || !L->dinv || !L->r || !L->z || !L->t) {
      mg_free(lev, nlev);
      return -1;
    }
  }
#pragma omp parallel for schedule(static)
  for (long long c = 0; c < (long long)n0; ++c) {
    const double* e = expr + c * 8;
    lev[0].kx[c] = -sgn * e[2];
    lev[0].ky[c] = -sgn * e[4];
    lev[0].kz[c] = -sgn * e[6];
    if (has_m) {
      const double rs = e[0] + e[1] + e[2] + e[3] + e[4] + e[5] + e[6];
      lev[0].m[c] = fabs(rs) > 1e-12 * fabs(e[0]) ? sgn * rs : 0.0;
    }
    if (coef16) {
      lev[0].kx[c] = bf16_round(lev[0].kx[c]);
      lev[0].ky[c] = bf16_round(lev[0].ky[c]);
      lev[0].kz[c] = bf16_round(lev[0].kz[c]);
      if (has_m) lev[0].m[c] = (double)(float)lev[0].m[c];
    }
  }
  for (int l = 1; l < nlev; ++l) {
    const mglev_t* F = &lev[l - 1];
    mglev_t* C = &lev[l];
#pragma omp parallel for schedule(static)
    for (int z = 0; z < C->nz; ++z)
      for (int y = 0; y < C->ny; ++y)
        for (int x = 0; x < C->nx; ++x) {
          const grid_t g = {F->nx, F->ny, F->nz};
          double sx = 0, sy = 0, sz = 0, sm = 0;
          for (int a = 0; a < 2; ++a)
            for (int bb = 0; bb < 2; ++bb) {
              sx += F->kx[idx(&g, 2 * x + 1, 2 * y + a, 2 * z + bb)];
              sy += F->ky[idx(&g, 2 * x + a, 2 * y + 1, 2 * z + bb)];
              sz += F->kz[idx(&g, 2 * x + a, 2 * y + bb, 2 * z + 1)];
            }
          const size_t c = (size_t)x + (size_t)C->nx * ((size_t)y + (size_t)C->ny * z);
          if (has_m) {
            float smf = 0;
            for (int d = 0; d < 8; ++d) {
              const double v =
                  F->m[idx(&g, 2 * x + (d & 1), 2 * y + ((d >> 1) & 1), 2 * z + (d >> 2))];
              sm += v;
              smf += (float)v;
            }
            C->m[c] = coef16 ? (double)(0.5f * smf) : 0.5 * sm;
          }
          C->kx[c] = coef16 ? bf16_round(0.5 * sx) : 0.5 * sx;
          C->ky[c] = coef16 ? bf16_round(0.5 * sy) : 0.5 * sy;
          C->kz[c] = coef16 ? bf16_round(0.5 * sz) : 0.5 * sz;
        }
  }
  for (int l = 0; l < nlev; ++l) {
    mglev_t* L = &lev[l];
    const grid_t g = {L->nx, L->ny, L->nz};
#pragma omp parallel for schedule(static)
    for (int z = 0; z < L->nz; ++z)
      for (int y = 0; y < L->ny; ++y)
        for (int x = 0; x < L->nx; ++x) {
          const size_t c = idx(&g, x, y, z);
          double d = L->kx[c] + L->kx[idx(&g, wrap(x - 1, L->nx), y, z)] + L->ky[c] +
                     L->ky[idx(&g, x, wrap(y - 1, L->ny), z)] + L->kz[c] +
                     L->kz[idx(&g, x, y, wrap(z - 1, L->nz))];
          if (L->m) d += L->m[c];
          if (coef16) {
            const float df = (float)d;
            L->dinv[c] = df != 0 ? bf16_round((double)(1.0f / df)) : 0.0;
          } else {
            L->dinv[c] = d != 0 ? 1.0 / d : 0.0;
          }
        }
  }
  return nlev;
}

/* z = M r for one right-hand side (tests: symmetry of M, action on modes);
 * coef16 as in mg_setup; returns nlev or -1/-2. */
int orc_mg_apply(const double* expr, int nx, int ny, int nz, int sweeps, int coef16, int w,
                 const double* r, double* z) {
  mglev_t lev[12];
  const int nlev = mg_setup(expr, nx, ny, nz, coef16, lev);
  if (nlev < 0) return nlev;
  memcpy(lev[0].r, r, lev[0].n * sizeof(double));
  mg_vcycle(lev, 0, nlev - 1, sweeps, w);
  memcpy(z, lev[0].z, lev[0].n * sizeof(double));
  mg_free(lev, nlev);
  return nlev;
}

int orc_mgpcg(const double* expr, int nx, int ny, int nz, const double* x0, double* x,
              double tol, int maxiter, int sweeps, int w, int remove_mean, double* relres_out,
              double* setup_s) {
  const double t0 = wall_seconds();
  mglev_t lev[12];
  const int nlev = mg_setup(expr, nx, ny, nz, 0, lev);
  if (nlev < 0) return nlev;
  if (setup_s) *setup_s = wall_seconds() - t0;
  mglev_t* L0 = &lev[0];
  const long long n = (long long)L0->n;
  double* r = (double*)malloc(n * sizeof(double));
  double* p = (double*)malloc(n * sizeof(double));
  double* q = (double*)malloc(n * sizeof(double));
  if (!r || !p || !q) {
    free(r); free(p); free(q);
    mg_free(lev, nlev);
    return -1;
  }
  double sgn = 1;
  for (long long c = 0; c < n; ++c)
    if (expr[c * 8] != 0) {
      sgn = expr[c * 8] < 0 ? -1 : 1;
      break;
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
    /* r = s * (b - A x) = -s * (A x + e7) */
#pragma omp parallel for schedule(static)
    for (int zz = 0; zz < nz; ++zz)
      for (int yy = 0; yy < ny; ++yy)
        for (int xx = 0; xx < nx; ++xx) {
          const grid_t g = {nx, ny, nz};
          const size_t c = idx(&g, xx, yy, zz);
          const double* e = expr + c * 8;
          size_t nb[6];
          neighbours(&g, xx, yy, zz, nb);
          double u = x[c] * e[0] + e[7];
          for (int k = 0; k < 6; ++k) u += x[nb[k]] * e[1 + k];
          r[c] = -sgn * u;
        }
    memcpy(L0->r, r, n * sizeof(double));
    mg_vcycle(lev, 0, nlev - 1, sweeps, w);
    double rz = 0, rr = 0;
#pragma omp parallel for reduction(+ : rz, rr) schedule(static)
    for (long long c = 0; c < n; ++c) {
      p[c] = L0->z[c];
      rz += r[c] * L0->z[c];
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
            const size_t c = (size_t)xx + (size_t)nx * ((size_t)yy + (size_t)ny * zz);
            const double u = mg_lap(L0, p, xx, yy, zz);
            q[c] = u;
            pq += p[c] * u;
          }
      const double alpha = rz / pq;
      rr = 0;
#pragma omp parallel for reduction(+ : rr) schedule(static)
      for (long long c = 0; c < n; ++c) {
        x[c] += alpha * p[c];
        r[c] -= alpha * q[c];
        L0->r[c] = r[c];
        rr += r[c] * r[c];
      }
      mg_vcycle(lev, 0, nlev - 1, sweeps, w);
      double rz_new = 0;
#pragma omp parallel for reduction(+ : rz_new) schedule(static)
      for (long long c = 0; c < n; ++c) rz_new += r[c] * L0->z[c];
      const double beta = rz_new / rz;
      rz = rz_new;
#pragma omp parallel for schedule(static)
      for (long long c = 0; c < n; ++c) p[c] = L0->z[c] + beta * p[c];
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
  free(r); free(p); free(q);
  mg_free(lev, nlev);
  return iter;
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
