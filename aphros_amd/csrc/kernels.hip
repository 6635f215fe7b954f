// HIP kernels of the aphpois engine (gfx950).
//
// Launch model: one launch covers every active system of an engine; a
// per-launch table maps each workgroup to (system, tile).  Scalars (alpha,
// beta, convergence flags) live in device memory; each reduction finishes in
// the last workgroup of its system (reduce.h), so a CG iteration needs no host
// round trip and a sequence of iterations can be captured in a hipGraph.  A
// converged system's workgroups return at entry.
//
// Stencil kernels march a (lx*VX) x (256/lx) tile of (x, y) columns through
// bz planes of z, keeping the z-1, z, z+1 values of the column in registers.
#include "kernels.h"
#include "reduce.h"

namespace aph {

template <typename T, int N>
using vec_t = T __attribute__((ext_vector_type(N)));

template <typename T, int VX>
__device__ __forceinline__ vec_t<double, VX> ldd(const T* __restrict__ p, long long i) {
  const vec_t<T, VX> v = *reinterpret_cast<const vec_t<T, VX>*>(p + i);
  return __builtin_convertvector(v, vec_t<double, VX>);
}

template <typename T, int VX>
__device__ __forceinline__ void std_(T* __restrict__ p, long long i, vec_t<double, VX> v) {
  *reinterpret_cast<vec_t<T, VX>*>(p + i) = __builtin_convertvector(v, vec_t<T, VX>);
}

// round a double vector through T (the value that gets stored)
template <typename T, int VX>
__device__ __forceinline__ vec_t<double, VX> rnd(vec_t<double, VX> v) {
  return __builtin_convertvector(__builtin_convertvector(v, vec_t<T, VX>), vec_t<double, VX>);
}

__device__ __forceinline__ int wrapm(int i, int n) { return i == 0 ? n - 1 : i - 1; }
__device__ __forceinline__ int wrapp(int i, int n) { return i == n - 1 ? 0 : i + 1; }

struct Tile {
  int x0, y, za, zb;
  bool active;
};

template <int VX>
__device__ __forceinline__ Tile tile_of(const Rec& rc, const SysDev& s, int lb) {
  Tile t;
  const int tx = lb % rc.tiles_x;
  const int t2 = lb / rc.tiles_x;
  const int ty = t2 % rc.tiles_y;
  const int tz = t2 / rc.tiles_y;
  const int rows = kBlock / rc.lx;
  const int ix = threadIdx.x % rc.lx, iy = threadIdx.x / rc.lx;
  t.x0 = (tx * rc.lx + ix) * VX;
  t.y = ty * rows + iy;
  t.za = tz * rc.bz;
  t.zb = min(t.za + rc.bz, s.nz);
  t.active = iy < rows && t.x0 < s.nx && t.y < s.ny;
  return t;
}

// ===========================================================================
// PCG (scaled, structure-preserving): q = S * L(k) * (S p) with
//   (L u)_c = sum_faces k_f (u_c - u_nb) + m_c u_c,
// where k are the plus-face conductances of each cell.  The p update
// p = r + beta * p_in is fused: the six neighbours' p are recomputed from r and
// p_in, and the new p goes to the other buffer p_out.
// ===========================================================================
template <typename T, int VX>
__global__ __launch_bounds__(kBlock) void k_pcg_spmv(const SysDev* __restrict__ sds,
                                                     const Rec* __restrict__ recs,
                                                     const int* __restrict__ b2r) {
  using D = vec_t<double, VX>;
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const Rec rc = recs[b2r[b]];
  const SysDev& s = sds[rc.sys];
  Scalars* sc = s.sc;
  if (sc->done) return;
  const int lb = b - rc.blk0;
  const Tile t = tile_of<VX>(rc, s, lb);
  const int nx = s.nx, ny = s.ny, nz = s.nz;
  const int cur = sc->cur;
  const double beta = sc->beta;
  const T* __restrict__ r = (const T*)s.r;
  const T* __restrict__ pin = (const T*)s.p[cur];
  T* __restrict__ pout = (T*)s.p[cur ^ 1];
  T* __restrict__ q = (T*)s.q;
  const T* __restrict__ kx = (const T*)s.kx;
  const T* __restrict__ ky = (const T*)s.ky;
  const T* __restrict__ kz = (const T*)s.kz;
  const T* __restrict__ S = (const T*)s.S;
  const T* __restrict__ mm = (const T*)s.m;
  const int has_m = s.has_m;
  double acc[1] = {0.0};
  if (t.active) {
    const long long pl = (long long)nx * ny;
    const long long row = (long long)t.y * nx;
    const long long rowm = (long long)wrapm(t.y, ny) * nx;
    const long long rowp = (long long)wrapp(t.y, ny) * nx;
    const int xm = t.x0 == 0 ? nx - 1 : t.x0 - 1;
    const int xp = t.x0 + VX == nx ? 0 : t.x0 + VX;
    auto pnew = [&](long long i) -> D {
      return rnd<T, VX>(ldd<T, VX>(r, i) + beta * ldd<T, VX>(pin, i));
    };
    auto pnew1 = [&](long long i) -> double {
      return (double)(T)((double)r[i] + beta * (double)pin[i]);
    };
    // prologue: plane za-1 and za
    const int zm0 = t.za == 0 ? nz - 1 : t.za - 1;
    long long i = (long long)zm0 * pl + row + t.x0;
    D um = ldd<T, VX>(S, i) * pnew(i);
    D kzm = ldd<T, VX>(kz, i);
    i = (long long)t.za * pl + row + t.x0;
    D pc = pnew(i);
    D Sc = ldd<T, VX>(S, i);
    D uc = Sc * pc;
    double dot = 0.0;
    for (int z = t.za; z < t.zb; ++z) {
      const long long base = (long long)z * pl;
      const long long ic = base + row + t.x0;
      const long long in = (long long)wrapp(z, nz) * pl + row + t.x0;
      const D pn = pnew(in);
      const D Sn = ldd<T, VX>(S, in);
      const D un = Sn * pn;
      const long long iym = base + rowm + t.x0, iyp = base + rowp + t.x0;
      const D uym = ldd<T, VX>(S, iym) * pnew(iym);
      const D uyp = ldd<T, VX>(S, iyp) * pnew(iyp);
      const long long ixm = base + row + xm, ixp = base + row + xp;
      const double uxm = (double)S[ixm] * pnew1(ixm);
      const double uxp = (double)S[ixp] * pnew1(ixp);
      const D kxc = ldd<T, VX>(kx, ic);
      const double kxm = (double)kx[ixm];
      const D kyc = ldd<T, VX>(ky, ic);
      const D kym = ldd<T, VX>(ky, iym);
      const D kzc = ldd<T, VX>(kz, ic);
      D l = kyc * (uc - uyp) + kym * (uc - uym) + kzc * (uc - un) + kzm * (uc - um);
#pragma unroll
      for (int v = 0; v < VX; ++v) {
        const double up = v + 1 < VX ? uc[(v + 1) % VX] : uxp;
        const double uw = v > 0 ? uc[(v + VX - 1) % VX] : uxm;
        const double kw = v > 0 ? kxc[(v + VX - 1) % VX] : kxm;
        l[v] += kxc[v] * (uc[v] - up) + kw * (uc[v] - uw);
      }
      if (has_m) l += ldd<T, VX>(mm, ic) * uc;
      const D qv = rnd<T, VX>(Sc * l);
      std_<T, VX>(pout, ic, pc);
      std_<T, VX>(q, ic, qv);
#pragma unroll
      for (int v = 0; v < VX; ++v) dot += pc[v] * qv[v];
      um = uc;
      uc = un;
      pc = pn;
      Sc = Sn;
      kzm = kzc;
    }
    acc[0] = dot;
  }
  if (grid_reduce<1, 0u>(acc, s.partials, rc.nblk, lb, s.counter)) {
    if (threadIdx.x == 0) {
      const double pq = acc[0];
      if (!(pq != 0.0) || !isfinite(pq)) {
        sc->status = 6;
        sc->done = 1;
      } else {
        sc->alpha = sc->rho / pq;
      }
      sc->cur = cur ^ 1;
    }
  }
}

// x^ += alpha p ; r -= alpha q ; (r,r).  Decides whether a reliable update
// (FP64 true residual) follows.
template <typename T, int VX>
__global__ __launch_bounds__(kBlock) void k_pcg_update(const SysDev* __restrict__ sds,
                                                       const Rec* __restrict__ recs,
                                                       const int* __restrict__ b2r) {
  using D = vec_t<double, VX>;
  const int b = blockIdx.x;
  const Rec rc = recs[b2r[b]];
  const SysDev& s = sds[rc.sys];
  Scalars* sc = s.sc;
  if (sc->done) return;
  const int lb = b - rc.blk0;
  const double alpha = sc->alpha;
  const T* __restrict__ p = (const T*)s.p[sc->cur];
  const T* __restrict__ q = (const T*)s.q;
  T* __restrict__ r = (T*)s.r;
  T* __restrict__ xh = (T*)s.xh;
  double acc[1] = {0.0};
  const long long stride = (long long)rc.nblk * kBlock * VX;
  for (long long i = ((long long)lb * kBlock + threadIdx.x) * VX; i < s.n; i += stride) {
    const D pv = ldd<T, VX>(p, i);
    const D xv = rnd<T, VX>(ldd<T, VX>(xh, i) + alpha * pv);
    const D rv = rnd<T, VX>(ldd<T, VX>(r, i) - alpha * ldd<T, VX>(q, i));
    std_<T, VX>(xh, i, xv);
    std_<T, VX>(r, i, rv);
#pragma unroll
    for (int v = 0; v < VX; ++v) acc[0] += rv[v] * rv[v];
  }
  if (grid_reduce<1, 0u>(acc, s.partials, rc.nblk, lb, s.counter)) {
    if (threadIdx.x == 0) {
      const double rn = acc[0];
      sc->iter += 1;
      const double est = sc->gamma * sqrt(rn);
      const bool trig = !isfinite(rn) || sqrt(rn) < sc->delta * sqrt(sc->rmax2) ||
                        est <= 0.9 * sc->tol || sc->iter >= sc->maxiter;
      if (trig) {
        sc->ru = 1;
        sc->rho_pend = rn;
      } else {
        sc->beta = rn / sc->rho;
        sc->rho = rn;
      }
    }
  }
}

// Reliable update, part 1: x64 = (x64 - mean_pending) + S * x^ ; x^ = 0 ;
// mean of the new x64 (nullspace removal).
template <typename T, int VX>
__global__ __launch_bounds__(kBlock) void k_ru_accum(const SysDev* __restrict__ sds,
                                                     const Rec* __restrict__ recs,
                                                     const int* __restrict__ b2r) {
  using D = vec_t<double, VX>;
  const int b = blockIdx.x;
  const Rec rc = recs[b2r[b]];
  const SysDev& s = sds[rc.sys];
  Scalars* sc = s.sc;
  if (sc->done || !sc->ru) return;
  const int lb = b - rc.blk0;
  const double mp = sc->mean_pending;
  const T* __restrict__ S = (const T*)s.S;
  T* __restrict__ xh = (T*)s.xh;
  double* __restrict__ x = s.x64;
  double acc[1] = {0.0};
  const long long stride = (long long)rc.nblk * kBlock * VX;
  for (long long i = ((long long)lb * kBlock + threadIdx.x) * VX; i < s.n; i += stride) {
    const D xv = (*reinterpret_cast<const D*>(x + i) - mp) + ldd<T, VX>(S, i) * ldd<T, VX>(xh, i);
    *reinterpret_cast<D*>(x + i) = xv;
    std_<T, VX>(xh, i, (D)(0.0));
#pragma unroll
    for (int v = 0; v < VX; ++v) acc[0] += xv[v];
  }
  if (grid_reduce<1, 0u>(acc, s.partials, rc.nblk, lb, s.counter)) {
    if (threadIdx.x == 0) {
      sc->mean_pending = sc->singular ? acc[0] / (double)s.n : 0.0;
    }
  }
}

// Reliable update, part 2: r64 = b - A (x64 - mean) in FP64 with the user's
// symmetric coefficients; r^ = sign * S * r64; stop test on the true residual;
// beta from the replaced residual.
template <typename T, int VX>
__global__ __launch_bounds__(kBlock) void k_ru_resid(const SysDev* __restrict__ sds,
                                                     const Rec* __restrict__ recs,
                                                     const int* __restrict__ b2r) {
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const Rec rc = recs[b2r[b]];
  const SysDev& s = sds[rc.sys];
  Scalars* sc = s.sc;
  if (sc->done || !sc->ru) return;
  const int lb = b - rc.blk0;
  const Tile t = tile_of<VX>(rc, s, lb);
  const int nx = s.nx, ny = s.ny, nz = s.nz;
  const double mp = sc->mean_pending;
  const double sgn = s.neg ? -1.0 : 1.0;
  const double* __restrict__ x = s.x64;
  const double* __restrict__ bb = s.b64;
  const double* __restrict__ e0 = s.s4[0];
  const double* __restrict__ ex = s.s4[1];
  const double* __restrict__ ey = s.s4[2];
  const double* __restrict__ ez = s.s4[3];
  const T* __restrict__ S = (const T*)s.S;
  T* __restrict__ r = (T*)s.r;
  double acc[3] = {0.0, 0.0, 0.0};
  if (t.active) {
    const long long pl = (long long)nx * ny;
    const long long row = (long long)t.y * nx;
    const long long rowm = (long long)wrapm(t.y, ny) * nx;
    const long long rowp = (long long)wrapp(t.y, ny) * nx;
    for (int z = t.za; z < t.zb; ++z) {
      const long long base = (long long)z * pl;
      const long long bm = (long long)wrapm(z, nz) * pl, bp = (long long)wrapp(z, nz) * pl;
      for (int v = 0; v < VX; ++v) {
        const int xx = t.x0 + v;
        const int xw = wrapm(xx, nx), xe = wrapp(xx, nx);
        const long long c = base + row + xx;
        const double xc = x[c] - mp;
        double ax = e0[c] * xc;
        ax += ex[base + row + xw] * (x[base + row + xw] - mp);
        ax += ex[c] * (x[base + row + xe] - mp);
        ax += ey[base + rowm + xx] * (x[base + rowm + xx] - mp);
        ax += ey[c] * (x[base + rowp + xx] - mp);
        ax += ez[bm + row + xx] * (x[bm + row + xx] - mp);
        ax += ez[c] * (x[bp + row + xx] - mp);
        const double r64 = bb[c] - ax;
        const T rh = (T)(sgn * (double)S[c] * r64);
        r[c] = rh;
        acc[0] += (double)rh * (double)rh;
        acc[1] += r64 * r64;
        acc[2] = fmax(acc[2], fabs(r64));
      }
    }
  }
  if (grid_reduce<3, 4u>(acc, s.partials, rc.nblk, lb, s.counter)) {
    if (threadIdx.x == 0) {
      const double rh2 = acc[0], rr = acc[1], mx = acc[2];
      const double relres = sc->bnorm2 > 0 ? sqrt(rr / sc->bnorm2) : sqrt(rr);
      const double l2 = sqrt(rr / sc->h3), lmax = mx / sc->h3;
      const double metric = sc->resmode == 2 ? relres : (sc->resmode == 1 ? lmax : l2);
      const bool conv = sc->resmode == 2 ? metric <= sc->tol : metric < sc->tol;
      sc->relres = relres;
      sc->res_l2 = l2;
      sc->res_max = lmax;
      sc->residual = metric;
      sc->nru += 1;
      if (!isfinite(rr) || !isfinite(rh2)) {
        sc->status = 6;
        sc->done = 1;
      } else if ((conv && sc->iter >= sc->miniter) || sc->iter >= sc->maxiter) {
        sc->converged = conv ? 1 : 0;
        sc->done = 1;
      } else {
        sc->beta = sc->first ? 0.0 : rh2 / sc->rho;
        sc->rho = rh2;
        sc->rmax2 = rh2;
        sc->gamma = rh2 > 0 ? metric / sqrt(rh2) : 1.0;
        sc->first = 0;
      }
      sc->ru = 0;
    }
  }
}

// ===========================================================================
// Reference-mode CG (SolverConjugate, src/linear/linear.ipp:24-129) on the
// general 7-coefficient Expr in FP64.  Per-cell arithmetic keeps the
// reference's operation order and no FMA contraction.
// ===========================================================================
#pragma clang fp contract(off)

__global__ __launch_bounds__(kBlock) void k_ref_init(const SysDev* __restrict__ sds,
                                                     const Rec* __restrict__ recs,
                                                     const int* __restrict__ b2r) {
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const Rec rc = recs[b2r[b]];
  const SysDev& s = sds[rc.sys];
  Scalars* sc = s.sc;
  if (sc->done) return;
  const int lb = b - rc.blk0;
  const Tile t = tile_of<1>(rc, s, lb);
  const int nx = s.nx, ny = s.ny, nz = s.nz;
  const double* __restrict__ u = s.u[0];
  double acc[1] = {0.0};
  if (t.active) {
    const long long pl = (long long)nx * ny;
    const long long row = (long long)t.y * nx;
    const long long rowm = (long long)wrapm(t.y, ny) * nx;
    const long long rowp = (long long)wrapp(t.y, ny) * nx;
    const int xx = t.x0;
    const int xw = wrapm(xx, nx), xe = wrapp(xx, nx);
    for (int z = t.za; z < t.zb; ++z) {
      const long long base = (long long)z * pl;
      const long long c = base + row + xx;
      double v = u[c] * s.f7[0][c] + s.f7[7][c];
      v += u[base + row + xw] * s.f7[1][c];
      v += u[base + row + xe] * s.f7[2][c];
      v += u[base + rowm + xx] * s.f7[3][c];
      v += u[base + rowp + xx] * s.f7[4][c];
      v += u[(long long)wrapm(z, nz) * pl + row + xx] * s.f7[5][c];
      v += u[(long long)wrapp(z, nz) * pl + row + xx] * s.f7[6][c];
      const double rv = -v;
      s.rr[c] = rv;
      s.pp[0][c] = rv;
      acc[0] += rv * rv;
    }
  }
  if (grid_reduce<1, 0u>(acc, s.partials, rc.nblk, lb, s.counter)) {
    if (threadIdx.x == 0) {
      sc->rho = acc[0];
      sc->cur = 0;
      sc->first = 1;
    }
  }
}

// p = r + beta p (skipped on the first iteration, where p = r) fused with
// q = A p and (p, q).
__global__ __launch_bounds__(kBlock) void k_ref_spmv(const SysDev* __restrict__ sds,
                                                     const Rec* __restrict__ recs,
                                                     const int* __restrict__ b2r) {
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const Rec rc = recs[b2r[b]];
  const SysDev& s = sds[rc.sys];
  Scalars* sc = s.sc;
  if (sc->done) return;
  const int lb = b - rc.blk0;
  const Tile t = tile_of<1>(rc, s, lb);
  const int nx = s.nx, ny = s.ny, nz = s.nz;
  const int cur = sc->cur;
  const int first = sc->first;
  const double beta = sc->beta;
  const double* __restrict__ r = s.rr;
  const double* __restrict__ pin = s.pp[cur];
  double* __restrict__ pout = s.pp[cur ^ 1];
  double acc[1] = {0.0};
  if (t.active) {
    auto pn = [&](long long i) -> double { return first ? pin[i] : r[i] + beta * pin[i]; };
    const long long pl = (long long)nx * ny;
    const long long row = (long long)t.y * nx;
    const long long rowm = (long long)wrapm(t.y, ny) * nx;
    const long long rowp = (long long)wrapp(t.y, ny) * nx;
    const int xx = t.x0;
    const int xw = wrapm(xx, nx), xe = wrapp(xx, nx);
    for (int z = t.za; z < t.zb; ++z) {
      const long long base = (long long)z * pl;
      const long long c = base + row + xx;
      const double pc = pn(c);
      double v = pc * s.f7[0][c];
      v += pn(base + row + xw) * s.f7[1][c];
      v += pn(base + row + xe) * s.f7[2][c];
      v += pn(base + rowm + xx) * s.f7[3][c];
      v += pn(base + rowp + xx) * s.f7[4][c];
      v += pn((long long)wrapm(z, nz) * pl + row + xx) * s.f7[5][c];
      v += pn((long long)wrapp(z, nz) * pl + row + xx) * s.f7[6][c];
      pout[c] = pc;
      s.qq[c] = v;
      acc[0] += pc * v;
    }
  }
  if (grid_reduce<1, 0u>(acc, s.partials, rc.nblk, lb, s.counter)) {
    if (threadIdx.x == 0) {
      sc->alpha = sc->rho / (acc[0] + 1e-100);
      sc->cur = cur ^ 1;
    }
  }
}

// u += alpha p ; r -= alpha q ; (r,r) and max|r| ; stop rule and beta
// (src/linear/linear.ipp:83-114).
__global__ __launch_bounds__(kBlock) void k_ref_update(const SysDev* __restrict__ sds,
                                                       const Rec* __restrict__ recs,
                                                       const int* __restrict__ b2r) {
  const int b = blockIdx.x;
  const Rec rc = recs[b2r[b]];
  const SysDev& s = sds[rc.sys];
  Scalars* sc = s.sc;
  if (sc->done) return;
  const int lb = b - rc.blk0;
  const double alpha = sc->alpha;
  const double* __restrict__ p = s.pp[sc->cur];
  double* __restrict__ u = s.u[0];
  double* __restrict__ r = s.rr;
  const double* __restrict__ q = s.qq;
  double acc[2] = {0.0, 0.0};
  for (long long i = (long long)lb * kBlock + threadIdx.x; i < s.n;
       i += (long long)rc.nblk * kBlock) {
    u[i] += alpha * p[i];
    const double rv = r[i] - alpha * q[i];
    r[i] = rv;
    acc[0] += rv * rv;
    acc[1] = fmax(acc[1], fabs(rv));
  }
  if (grid_reduce<2, 2u>(acc, s.partials, rc.nblk, lb, s.counter)) {
    if (threadIdx.x == 0) {
      const double dot_r = acc[0];
      const double res = sc->resmode == 1 ? acc[1] / sc->h3 : sqrt(dot_r / sc->h3);
      sc->residual = res;
      sc->iter += 1;
      sc->beta = dot_r / (sc->rho + 1e-100);
      sc->rho = dot_r;
      sc->first = 0;
      if (sc->iter >= sc->miniter && (sc->iter > sc->maxiter || res < sc->tol)) {
        sc->converged = res < sc->tol ? 1 : 0;
        sc->done = 1;
      }
    }
  }
}

// Jacobi sweep (src/linear/linear.ipp:179-201): u_new = -(e7 + sum e_q u_nb)/e0.
__global__ __launch_bounds__(kBlock) void k_jacobi(const SysDev* __restrict__ sds,
                                                   const Rec* __restrict__ recs,
                                                   const int* __restrict__ b2r) {
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const Rec rc = recs[b2r[b]];
  const SysDev& s = sds[rc.sys];
  Scalars* sc = s.sc;
  if (sc->done) return;
  const int lb = b - rc.blk0;
  const Tile t = tile_of<1>(rc, s, lb);
  const int nx = s.nx, ny = s.ny, nz = s.nz;
  const int cur = sc->cur;
  const double* __restrict__ u = s.u[cur];
  double* __restrict__ un = s.u[cur ^ 1];
  double acc[1] = {0.0};
  if (t.active) {
    const long long pl = (long long)nx * ny;
    const long long row = (long long)t.y * nx;
    const long long rowm = (long long)wrapm(t.y, ny) * nx;
    const long long rowp = (long long)wrapp(t.y, ny) * nx;
    const int xx = t.x0;
    const int xw = wrapm(xx, nx), xe = wrapp(xx, nx);
    for (int z = t.za; z < t.zb; ++z) {
      const long long base = (long long)z * pl;
      const long long c = base + row + xx;
      double nd = s.f7[7][c];
      nd += u[base + row + xw] * s.f7[1][c];
      nd += u[base + row + xe] * s.f7[2][c];
      nd += u[base + rowm + xx] * s.f7[3][c];
      nd += u[base + rowp + xx] * s.f7[4][c];
      nd += u[(long long)wrapm(z, nz) * pl + row + xx] * s.f7[5][c];
      nd += u[(long long)wrapp(z, nz) * pl + row + xx] * s.f7[6][c];
      const double v = -nd / s.f7[0][c];
      un[c] = v;
      acc[0] = fmax(acc[0], fabs(v - u[c]));
    }
  }
  if (grid_reduce<1, 1u>(acc, s.partials, rc.nblk, lb, s.counter)) {
    if (threadIdx.x == 0) {
      const double res = acc[0];
      sc->residual = res;
      sc->iter += 1;
      sc->cur = cur ^ 1;
      if (sc->iter >= sc->miniter && (sc->iter > sc->maxiter || res < sc->tol)) {
        sc->converged = res < sc->tol ? 1 : 0;
        sc->done = 1;
      }
    }
  }
}

// True residual of a general Expr: (A x + e7), reported after reference-mode
// solves.  acc: sum r^2, sum e7^2, max |r|.
__global__ __launch_bounds__(kBlock) void k_true_res_f7(const SysDev* __restrict__ sds,
                                                        const Rec* __restrict__ recs,
                                                        const int* __restrict__ b2r) {
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const Rec rc = recs[b2r[b]];
  const SysDev& s = sds[rc.sys];
  Scalars* sc = s.sc;
  const int lb = b - rc.blk0;
  const Tile t = tile_of<1>(rc, s, lb);
  const int nx = s.nx, ny = s.ny, nz = s.nz;
  const double* __restrict__ u = sc->method == kJacobi ? s.u[sc->cur] : s.u[0];
  double acc[3] = {0.0, 0.0, 0.0};
  if (t.active) {
    const long long pl = (long long)nx * ny;
    const long long row = (long long)t.y * nx;
    const long long rowm = (long long)wrapm(t.y, ny) * nx;
    const long long rowp = (long long)wrapp(t.y, ny) * nx;
    const int xx = t.x0;
    const int xw = wrapm(xx, nx), xe = wrapp(xx, nx);
    for (int z = t.za; z < t.zb; ++z) {
      const long long base = (long long)z * pl;
      const long long c = base + row + xx;
      double v = u[c] * s.f7[0][c] + s.f7[7][c];
      v += u[base + row + xw] * s.f7[1][c];
      v += u[base + row + xe] * s.f7[2][c];
      v += u[base + rowm + xx] * s.f7[3][c];
      v += u[base + rowp + xx] * s.f7[4][c];
      v += u[(long long)wrapm(z, nz) * pl + row + xx] * s.f7[5][c];
      v += u[(long long)wrapp(z, nz) * pl + row + xx] * s.f7[6][c];
      acc[0] += v * v;
      acc[1] += s.f7[7][c] * s.f7[7][c];
      acc[2] = fmax(acc[2], fabs(v));
    }
  }
  if (grid_reduce<3, 4u>(acc, s.partials, rc.nblk, lb, s.counter)) {
    if (threadIdx.x == 0) {
      sc->relres = acc[1] > 0 ? sqrt(acc[0] / acc[1]) : sqrt(acc[0]);
      sc->res_l2 = sqrt(acc[0] / sc->h3);
      sc->res_max = acc[2] / sc->h3;
    }
  }
}

#pragma clang fp contract(on)

// ===========================================================================
// System conversion (set_system): AoS Expr with strides -> SoA layouts.
// ===========================================================================

// flags: 1 non-finite, 2 asymmetric, 4 zero diagonal, 16 positive diagonal
// seen, 32 negative diagonal seen, 64 non-zero row sum (mass term)
__global__ __launch_bounds__(kBlock) void k_check(const double* __restrict__ aos, long long start,
                                                  long long ly, long long lz, int nx, int ny,
                                                  int nz, unsigned* __restrict__ flags) {
  __shared__ unsigned sflags;
  if (threadIdx.x == 0) sflags = 0;
  __syncthreads();
  const long long n = (long long)nx * ny * nz;
  unsigned f = 0;
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    const int x = (int)(i % nx);
    const long long t = i / nx;
    const int y = (int)(t % ny);
    const int z = (int)(t / ny);
    const double* e = aos + 8 * (start + z * lz + y * ly + x);
    const double* ew = aos + 8 * (start + z * lz + y * ly + wrapm(x, nx));
    const double* es = aos + 8 * (start + z * lz + wrapm(y, ny) * ly + x);
    const double* eb = aos + 8 * (start + wrapm(z, nz) * lz + y * ly + x);
    for (int k = 0; k < 8; ++k) {
      if (!isfinite(e[k])) f |= 1u;
    }
    if (e[1] != ew[2] || e[3] != es[4] || e[5] != eb[6]) f |= 2u;
    if (e[0] == 0.0) f |= 4u;
    if (e[0] > 0) f |= 16u;
    if (e[0] < 0) f |= 32u;
    const double rs = e[0] + e[1] + e[2] + e[3] + e[4] + e[5] + e[6];
    if (fabs(rs) > 1e-12 * fabs(e[0])) f |= 64u;
  }
  if (f) atomicOr(&sflags, f);
  __syncthreads();
  if (threadIdx.x == 0 && sflags) atomicOr(flags, sflags);
}

__global__ __launch_bounds__(kBlock) void k_convert_f7(const double* __restrict__ aos,
                                                       long long start, long long ly,
                                                       long long lz, int nx, int ny, int nz,
                                                       double* __restrict__ f7, int coefs_only) {
  const long long n = (long long)nx * ny * nz;
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    const int x = (int)(i % nx);
    const long long t = i / nx;
    const int y = (int)(t % ny);
    const int z = (int)(t / ny);
    const double* e = aos + 8 * (start + z * lz + y * ly + x);
    const int kend = coefs_only ? 7 : 8;
    for (int k = 0; k < kend; ++k) f7[(long long)k * n + i] = e[k];
  }
}

// PCG storage: sign s of the diagonal, conductances k = -s * e_plus, scale
// S = 1/sqrt(|e0|), mass m = s * (row sum), FP64 copies for the true residual
// and b = -e7.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_convert_pcg(
    const double* __restrict__ aos, long long start, long long ly, long long lz, int nx, int ny,
    int nz, double sgn, T* __restrict__ kx, T* __restrict__ ky, T* __restrict__ kz,
    T* __restrict__ S, T* __restrict__ m, double* __restrict__ s4, double* __restrict__ b,
    int coefs_only) {
  const long long n = (long long)nx * ny * nz;
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    const int x = (int)(i % nx);
    const long long t = i / nx;
    const int y = (int)(t % ny);
    const int z = (int)(t / ny);
    const double* e = aos + 8 * (start + z * lz + y * ly + x);
    if (!coefs_only) b[i] = -e[7];
    if (coefs_only == 2) continue;  // rhs only
    kx[i] = (T)(-sgn * e[2]);
    ky[i] = (T)(-sgn * e[4]);
    kz[i] = (T)(-sgn * e[6]);
    S[i] = (T)(1.0 / sqrt(fabs(e[0])));
    if (m) {
      const double rs = e[0] + e[1] + e[2] + e[3] + e[4] + e[5] + e[6];
      m[i] = fabs(rs) > 1e-12 * fabs(e[0]) ? (T)(sgn * rs) : (T)0;
    }
    s4[i] = e[0];
    s4[n + i] = e[2];
    s4[2 * n + i] = e[4];
    s4[3 * n + i] = e[6];
  }
}

// constant term only, from a strided host-layout copy: dst = scale * e7
__global__ __launch_bounds__(kBlock) void k_convert_rhs(const double* __restrict__ src,
                                                        long long start, long long ly,
                                                        long long lz, int nx, int ny, int nz,
                                                        double scale, double* __restrict__ dst) {
  const long long n = (long long)nx * ny * nz;
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    const int x = (int)(i % nx);
    const long long t = i / nx;
    const int y = (int)(t % ny);
    const int z = (int)(t / ny);
    dst[i] = scale * src[start + z * lz + y * ly + x];
  }
}

// gather/scatter of a scalar field between strided and dense layouts
__global__ __launch_bounds__(kBlock) void k_gather(const double* __restrict__ src, long long start,
                                                   long long ly, long long lz, int nx, int ny,
                                                   int nz, double* __restrict__ dst, double sub) {
  const long long n = (long long)nx * ny * nz;
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    const int x = (int)(i % nx);
    const long long t = i / nx;
    const int y = (int)(t % ny);
    const int z = (int)(t / ny);
    dst[i] = src[start + z * lz + y * ly + x] - sub;
  }
}

__global__ __launch_bounds__(kBlock) void k_scatter(const double* __restrict__ src, int nx,
                                                    int ny, int nz, double* __restrict__ dst,
                                                    long long start, long long ly, long long lz,
                                                    double sub) {
  const long long n = (long long)nx * ny * nz;
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    const int x = (int)(i % nx);
    const long long t = i / nx;
    const int y = (int)(t % ny);
    const int z = (int)(t / ny);
    dst[start + z * lz + y * ly + x] = src[i] - sub;
  }
}

// sum of squares of a dense FP64 array into out[0] (two-stage, deterministic)
__global__ __launch_bounds__(kBlock) void k_sumsq(const double* __restrict__ v, long long n,
                                                  double* __restrict__ partials,
                                                  unsigned* __restrict__ counter,
                                                  double* __restrict__ out) {
  double acc[1] = {0.0};
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    acc[0] += v[i] * v[i];
  }
  if (grid_reduce<1, 0u>(acc, partials, gridDim.x, blockIdx.x, counter)) {
    if (threadIdx.x == 0) *out = acc[0];
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill(T* __restrict__ p, long long n, T v) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    p[i] = v;
  }
}

// ===========================================================================
// launchers
// ===========================================================================

#define APH_LAUNCH(kern, grid, stream, ...) \
  hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), 0, stream, __VA_ARGS__)

hipError_t launch_iter(int method, int fp32, int vx, const LaunchSet& L, hipStream_t st) {
  if (method == kConjugate) {
    APH_LAUNCH(k_ref_spmv, L.nst, st, L.sds, L.rec_st, L.b2r_st);
    APH_LAUNCH(k_ref_update, L.nsm, st, L.sds, L.rec_sm, L.b2r_sm);
  } else if (method == kJacobi) {
    APH_LAUNCH(k_jacobi, L.nst, st, L.sds, L.rec_st, L.b2r_st);
  } else {
#define APH_PCG(T, VX)                                                  \
  do {                                                                  \
    APH_LAUNCH((k_pcg_spmv<T, VX>), L.nst, st, L.sds, L.rec_st, L.b2r_st);   \
    APH_LAUNCH((k_pcg_update<T, VX>), L.nsm, st, L.sds, L.rec_sm, L.b2r_sm); \
    APH_LAUNCH((k_ru_accum<T, VX>), L.nsm, st, L.sds, L.rec_sm, L.b2r_sm);   \
    APH_LAUNCH((k_ru_resid<T, 1>), L.nru, st, L.sds, L.rec_ru, L.b2r_ru);    \
  } while (0)
    if (fp32) {
      if (vx == 4) APH_PCG(float, 4);
      else if (vx == 2) APH_PCG(float, 2);
      else APH_PCG(float, 1);
    } else {
      if (vx == 2) APH_PCG(double, 2);
      else APH_PCG(double, 1);
    }
#undef APH_PCG
  }
  return hipGetLastError();
}

hipError_t launch_start(int method, int fp32, int vx, const LaunchSet& L, hipStream_t st) {
  if (method == kConjugate) {
    APH_LAUNCH(k_ref_init, L.nst, st, L.sds, L.rec_st, L.b2r_st);
  } else if (method == kPcg) {
#define APH_RU(T, VX)                                                  \
  do {                                                                 \
    APH_LAUNCH((k_ru_accum<T, VX>), L.nsm, st, L.sds, L.rec_sm, L.b2r_sm); \
    APH_LAUNCH((k_ru_resid<T, 1>), L.nru, st, L.sds, L.rec_ru, L.b2r_ru);  \
  } while (0)
    if (fp32) {
      if (vx == 4) APH_RU(float, 4);
      else if (vx == 2) APH_RU(float, 2);
      else APH_RU(float, 1);
    } else {
      if (vx == 2) APH_RU(double, 2);
      else APH_RU(double, 1);
    }
#undef APH_RU
  }
  return hipGetLastError();
}

hipError_t launch_true_res_f7(const LaunchSet& L, hipStream_t st) {
  APH_LAUNCH(k_true_res_f7, L.nru, st, L.sds, L.rec_ru, L.b2r_ru);
  return hipGetLastError();
}

static int grid_for(long long n) {
  long long g = (n + kBlock - 1) / kBlock;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  return (int)g;
}

hipError_t launch_check(const double* aos, long long start, long long ly, long long lz, int nx,
                        int ny, int nz, unsigned* flags, hipStream_t st) {
  APH_LAUNCH(k_check, grid_for((long long)nx * ny * nz), st, aos, start, ly, lz, nx, ny, nz,
             flags);
  return hipGetLastError();
}

hipError_t launch_convert_f7(const double* aos, long long start, long long ly, long long lz,
                             int nx, int ny, int nz, double* f7, int coefs_only, hipStream_t st) {
  APH_LAUNCH(k_convert_f7, grid_for((long long)nx * ny * nz), st, aos, start, ly, lz, nx, ny,
             nz, f7, coefs_only);
  return hipGetLastError();
}

hipError_t launch_convert_pcg(int fp32, const double* aos, long long start, long long ly,
                              long long lz, int nx, int ny, int nz, double sgn, void* kx,
                              void* ky, void* kz, void* S, void* m, double* s4, double* b,
                              int coefs_only, hipStream_t st) {
  const int g = grid_for((long long)nx * ny * nz);
  if (fp32) {
    APH_LAUNCH(k_convert_pcg<float>, g, st, aos, start, ly, lz, nx, ny, nz, sgn, (float*)kx,
               (float*)ky, (float*)kz, (float*)S, (float*)m, s4, b, coefs_only);
  } else {
    APH_LAUNCH(k_convert_pcg<double>, g, st, aos, start, ly, lz, nx, ny, nz, sgn, (double*)kx,
               (double*)ky, (double*)kz, (double*)S, (double*)m, s4, b, coefs_only);
  }
  return hipGetLastError();
}

hipError_t launch_convert_rhs(const double* src, long long start, long long ly, long long lz,
                              int nx, int ny, int nz, double scale, double* dst, hipStream_t st) {
  APH_LAUNCH(k_convert_rhs, grid_for((long long)nx * ny * nz), st, src, start, ly, lz, nx, ny,
             nz, scale, dst);
  return hipGetLastError();
}

hipError_t launch_gather(const double* src, long long start, long long ly, long long lz, int nx,
                         int ny, int nz, double* dst, double sub, hipStream_t st) {
  APH_LAUNCH(k_gather, grid_for((long long)nx * ny * nz), st, src, start, ly, lz, nx, ny, nz,
             dst, sub);
  return hipGetLastError();
}

hipError_t launch_scatter(const double* src, int nx, int ny, int nz, double* dst,
                          long long start, long long ly, long long lz, double sub,
                          hipStream_t st) {
  APH_LAUNCH(k_scatter, grid_for((long long)nx * ny * nz), st, src, nx, ny, nz, dst, start, ly,
             lz, sub);
  return hipGetLastError();
}

hipError_t launch_sumsq(const double* v, long long n, double* partials, unsigned* counter,
                        double* out, hipStream_t st) {
  int g = grid_for(n);
  if (g > 1024) g = 1024;
  APH_LAUNCH(k_sumsq, g, st, v, n, partials, counter, out);
  return hipGetLastError();
}

hipError_t launch_fill(void* p, long long n, int elem_bytes, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (elem_bytes == 8) {
    APH_LAUNCH(k_fill<double>, grid_for(n), st, (double*)p, n, 0.0);
  } else {
    APH_LAUNCH(k_fill<float>, grid_for(n), st, (float*)p, n, 0.0f);
  }
  return hipGetLastError();
}

}  // namespace aph
