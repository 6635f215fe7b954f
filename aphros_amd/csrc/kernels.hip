// HIP kernels of the aphpois engine (gfx950).
//
// Launch model: one launch covers every active system of an engine; a
// per-launch table maps each workgroup to (system, tile).  Scalars (alpha,
// beta, convergence flags) live in device memory; every reducing kernel writes
// per-workgroup partials and is followed by a finalize kernel with one
// workgroup per system (reduce.h), so a CG iteration needs no host round trip
// and a sequence of iterations can be captured in a hipGraph.  A converged
// system's workgroups return at entry.
//
// Stencil kernels march a (lx*VX) x (256/lx) tile of (x, y) columns through
// bz planes of z, keeping the z-1, z, z+1 values of the column in registers.
#include <type_traits>

#include "kernels.h"
#include "reduce.h"

namespace aph {

template <typename T, int N>
using vec_t = T __attribute__((ext_vector_type(N)));

// Global memory through pointers taken from the system tables: those are
// generic pointers, and accesses through them would compile to flat
// instructions (64-bit addresses in VGPRs, also counted by the LDS wait
// counter).  The helpers below address global memory explicitly; they are
// only used on global arrays.
template <typename T>
using gcp = const __attribute__((address_space(1))) T*;
template <typename T>
using gmp = __attribute__((address_space(1))) T*;

template <typename T, int VX, typename P>
__device__ __forceinline__ vec_t<double, VX> ldd(P p, long long i) {
  const vec_t<T, VX> v = *(gcp<vec_t<T, VX>>)(p + i);
  return __builtin_convertvector(v, vec_t<double, VX>);
}

template <typename T, int VX, typename P>
__device__ __forceinline__ void std_(P p, long long i, vec_t<double, VX> v) {
  *(gmp<vec_t<T, VX>>)(p + i) = __builtin_convertvector(v, vec_t<T, VX>);
}

template <typename T, int VX, typename P>
__device__ __forceinline__ vec_t<T, VX> ldv(P p, long long i) {
  return *(gcp<vec_t<T, VX>>)(p + i);
}

template <typename T, int VX, typename P>
__device__ __forceinline__ void stv(P p, long long i, vec_t<T, VX> v) {
  *(gmp<vec_t<T, VX>>)(p + i) = v;
}

template <typename T>
__device__ __forceinline__ T ld1(const T* __restrict__ p, long long i) {
  return *(gcp<T>)(p + i);
}

template <typename T>
__device__ __forceinline__ void st1(T* __restrict__ p, long long i, T v) {
  *(gmp<T>)(p + i) = v;
}

// A copy of a table entry that no kernel writes while the reader runs
// (system, level and launch tables), read through the constant address
// space: uniform scalar loads into SGPRs, so the pointers it holds need no
// VGPRs and the loads through them can take a scalar base.
//
// Only global memory may be read this way.  A generic pointer into LDS or
// into a thread's private memory (a __shared__ table, a by-value copy of a
// table entry) holds an aperture address that is not a global address;
// reinterpreted as a constant-space pointer it would read unmapped memory
// and fault.  So cload takes an explicitly global pointer: call sites state
// with gtab() that the table is a hipMalloc'ed device array, and an LDS or
// stack address does not convert to it implicitly.
template <typename T>
__device__ __forceinline__ gcp<T> gtab(const T* p) {
  return (gcp<T>)p;
}
template <typename T>
__device__ __forceinline__ T cload(gcp<T> p) {
  static_assert(sizeof(T) % 4 == 0 && std::is_trivially_copyable<T>::value, "dword copy");
  T v;
  const __attribute__((address_space(4))) unsigned* src =
      (const __attribute__((address_space(4))) unsigned*)p;
  unsigned* dst = reinterpret_cast<unsigned*>(&v);
#pragma unroll
  for (int k = 0; k < (int)(sizeof(T) / 4); ++k) dst[k] = src[k];
  return v;
}

// bfloat16 (the upper 16 bits of an IEEE float): the V-cycle coefficients of
// FP32 engines.  The V-cycle only has to be a fixed symmetric positive
// definite operator, so 8-bit mantissas in its coefficients cost nothing in
// convergence and halve their traffic.
__device__ __forceinline__ float bf2f(unsigned short b) {
  return __uint_as_float((unsigned)b << 16);
}
// round to nearest even (finite values)
__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

template <typename A, int VX, typename V>
__device__ __forceinline__ vec_t<A, VX> cva(V v) {
  return __builtin_convertvector(v, vec_t<A, VX>);
}

// Coefficients of a multigrid level: component c = 0..3 (kx, ky, kz, D^-1)
// of the VX cells from flat index i (i % VX == 0).  FP64: separate arrays;
// FP32: bfloat16 packed by groups of 4 cells (16 values per group: kx[4],
// ky[4], kz[4], dinv[4]), so a thread with 4 cells reads all four components
// of its cells with two 16-byte loads.
template <typename T>
struct MgCoefs {
  const T* a[4];
  __device__ explicit MgCoefs(const MgLevel& L)
      : a{(const T*)L.kx, (const T*)L.ky, (const T*)L.kz, (const T*)L.dinv} {}
  template <int VX>
  __device__ vec_t<T, VX> comp(int c, long long i) const {
    return ldv<T, VX>(a[c], i);
  }
  __device__ T comp1(int c, long long i) const { return a[c][i]; }
  __device__ T comp1g(int c, long long i) const { return ld1(a[c], i); }
  template <int VX>
  __device__ void own(long long i, vec_t<T, VX>& kx, vec_t<T, VX>& ky, vec_t<T, VX>& kz,
                      vec_t<T, VX>& d) const {
    kx = comp<VX>(0, i);
    ky = comp<VX>(1, i);
    kz = comp<VX>(2, i);
    d = comp<VX>(3, i);
  }
};

template <>
struct MgCoefs<float> {
  const unsigned short* kc;
  __device__ explicit MgCoefs(const MgLevel& L) : kc((const unsigned short*)L.kc) {}
  __device__ static long long at(int c, long long i) { return ((i >> 2) << 4) + 4 * c + (i & 3); }
  template <int VX>
  __device__ vec_t<float, VX> comp(int c, long long i) const {
    vec_t<float, VX> r;
    const unsigned short* p = kc + at(c, i);
    if constexpr (VX == 4) {
      const vec_t<unsigned, 2> w = *(gcp<vec_t<unsigned, 2>>)p;
      r[0] = __uint_as_float(w.x << 16);
      r[1] = __uint_as_float(w.x & 0xffff0000u);
      r[2] = __uint_as_float(w.y << 16);
      r[3] = __uint_as_float(w.y & 0xffff0000u);
    } else if constexpr (VX == 2) {
      const unsigned w = *(gcp<unsigned>)p;
      r[0] = __uint_as_float(w << 16);
      r[1] = __uint_as_float(w & 0xffff0000u);
    } else {
      r[0] = bf2f(ld1(p, 0));
    }
    return r;
  }
  __device__ float comp1(int c, long long i) const { return bf2f(kc[at(c, i)]); }
  __device__ float comp1g(int c, long long i) const { return bf2f(ld1(kc, at(c, i))); }
  template <int VX>
  __device__ void own(long long i, vec_t<float, VX>& kx, vec_t<float, VX>& ky,
                      vec_t<float, VX>& kz, vec_t<float, VX>& d) const {
    if constexpr (VX == 4) {
      const gcp<vec_t<unsigned, 4>> g = (gcp<vec_t<unsigned, 4>>)(kc + ((i >> 2) << 4));
      const vec_t<unsigned, 4> w0 = g[0], w1 = g[1];
      kx[0] = __uint_as_float(w0.x << 16);
      kx[1] = __uint_as_float(w0.x & 0xffff0000u);
      kx[2] = __uint_as_float(w0.y << 16);
      kx[3] = __uint_as_float(w0.y & 0xffff0000u);
      ky[0] = __uint_as_float(w0.z << 16);
      ky[1] = __uint_as_float(w0.z & 0xffff0000u);
      ky[2] = __uint_as_float(w0.w << 16);
      ky[3] = __uint_as_float(w0.w & 0xffff0000u);
      kz[0] = __uint_as_float(w1.x << 16);
      kz[1] = __uint_as_float(w1.x & 0xffff0000u);
      kz[2] = __uint_as_float(w1.y << 16);
      kz[3] = __uint_as_float(w1.y & 0xffff0000u);
      d[0] = __uint_as_float(w1.z << 16);
      d[1] = __uint_as_float(w1.z & 0xffff0000u);
      d[2] = __uint_as_float(w1.w << 16);
      d[3] = __uint_as_float(w1.w & 0xffff0000u);
    } else {
      kx = comp<VX>(0, i);
      ky = comp<VX>(1, i);
      kz = comp<VX>(2, i);
      d = comp<VX>(3, i);
    }
  }
};

// stores component c of cell i (setup kernels)
template <typename T>
__device__ __forceinline__ void st_coef(const MgLevel& L, int c, long long i, T v) {
  if constexpr (std::is_same<T, float>::value) {
    ((unsigned short*)L.kc)[MgCoefs<float>::at(c, i)] = f2bf(v);
  } else {
    const void* a[4] = {L.kx, L.ky, L.kz, L.dinv};
    ((T*)a[c])[i] = v;
  }
}

template <int VX, typename V>
__device__ __forceinline__ vec_t<double, VX> cvt(V v) {
  return __builtin_convertvector(v, vec_t<double, VX>);
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
//
// Each thread owns VX consecutive cells of one row of the tile and marches
// through z.  u = S * p_new of the own cells is computed once per plane and
// kept in registers for z-1, z, z+1; the plane's u of the whole tile is also
// put in LDS (double-buffered by plane) so that x and y neighbours inside the
// tile come from LDS.  Neighbours outside the tile (tile edges, periodic
// wrap) are recomputed from global memory by the edge threads.  u is formed
// in FP64 from the stored (rounded) p and S, so the differences u_c - u_nb are
// exact and q carries the rounding of its final store only.
//
// SC = false (multigrid mode): S is 1 and the direction is built from the
// preconditioned residual z: p = z + beta p_in, q = L p.
// XH: x^ += alpha p_in of the previous iteration, on the tile's own cells
// where p_in is read (planes za .. zb-1), instead of a pass of the update
// kernel; skipped when that iteration's reliable update folded alpha p into
// x64, and x^ = alpha p_in (no read) in the iteration after, x^ being stale
// since that update (xh_fold 1, 2).
template <typename T, int VX, bool SC, bool XH>
__global__ __launch_bounds__(kBlock) void k_pcg_spmv(const SysDev* __restrict__ sds,
                                                     const Rec* __restrict__ recs,
                                                     const int* __restrict__ b2r) {
  // arithmetic of the operator: FP64 for the Jacobi-scaled form, T with the
  // multigrid preconditioner (whose V-cycle absorbs FP32 rounding anyway)
  using A = typename std::conditional<SC, double, T>::type;
  using D = vec_t<A, VX>;
  __shared__ A su[2][kBlock * VX];
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const Rec rc = recs[b2r[b]];
  const SysDev& s = sds[rc.sys];
  Scalars* sc = s.sc;
  if (sc->done || sc->ru) return;  // ru: a reliable update is pending
  const int lb = b - rc.blk0;
  const Tile t = tile_of<VX>(rc, s, lb);
  const int nx = s.nx, ny = s.ny, nz = s.nz;
  const int cur = sc->cur;
  const double beta = sc->beta;
  const gcp<T> __restrict__ r = (gcp<T>)s.z;  // r (Jacobi-scaled) or z = M^-1 r
  const gcp<T> __restrict__ pin = (gcp<T>)s.p[cur];
  const gmp<T> __restrict__ pout = (gmp<T>)s.p[cur ^ 1];
  const gmp<T> __restrict__ q = (gmp<T>)s.q;
  const gcp<T> __restrict__ kx = (gcp<T>)s.kx;
  const gcp<T> __restrict__ ky = (gcp<T>)s.ky;
  const gcp<T> __restrict__ kz = (gcp<T>)s.kz;
  const gcp<T> __restrict__ S = (gcp<T>)s.S;
  const gcp<T> __restrict__ mm = (gcp<T>)s.m;
  const int has_m = s.has_m;
  const int lx = rc.lx;
  const int tx = lx * VX;                 // tile width in cells (LDS row)
  const int ix = threadIdx.x % lx, iy = threadIdx.x / lx;
  const int ty = kBlock / lx;
  const int lofs = iy * tx + ix * VX;     // own slot in the LDS plane
  // which neighbours are inside the tile (and the grid): others from global
  const bool xm_in = ix > 0;
  const bool xp_in = ix + 1 < lx && t.x0 + VX < nx;
  const bool ym_in = iy > 0;
  const bool yp_in = iy + 1 < ty && t.y + 1 < ny;
  const long long pl = (long long)nx * ny;
  const long long row = (long long)t.y * nx;
  const long long rowm = (long long)wrapm(t.y, ny) * nx;
  const long long rowp = (long long)wrapp(t.y, ny) * nx;
  const int xm = t.x0 == 0 ? nx - 1 : t.x0 - 1;
  const int xp = t.x0 + VX == nx ? 0 : t.x0 + VX;
  using V = vec_t<T, VX>;
  // stored value of p_new = r + beta * p_in (rounded to T)
  auto pnew_st = [&](long long i) -> V {
    return __builtin_convertvector(ldd<T, VX>(r, i) + beta * ldd<T, VX>(pin, i), V);
  };
  const int xs = XH ? sc->xh_fold : 1;  // 0 x^ += xa p_in, 1 none, 2 x^ = xa p_in
  const double xa = sc->alpha;
  const gmp<T> __restrict__ xh = (gmp<T>)s.xh;
  // pnew_st, and the x^ update on cells the tile owns
  auto pnew_own = [&](long long i) -> V {
    const vec_t<double, VX> po = ldd<T, VX>(pin, i);
    if constexpr (XH) {
      using D1 = vec_t<double, VX>;
      if (xs == 0) {
        if (xa != 0.0) std_<T, VX>(xh, i, ldd<T, VX>(xh, i) + xa * po);
      } else if (xs == 2) {
        std_<T, VX>(xh, i, (D1)(0.0) + xa * po);  // the sum with x^ = 0, to the bit
      }
    }
    return __builtin_convertvector(ldd<T, VX>(r, i) + beta * po, V);
  };
  auto sld = [&](long long i) -> V {
    if constexpr (SC) return ldv<T, VX>(S, i);
    return (T)1;
  };
  auto unew = [&](long long i) -> D {
    if constexpr (SC) return ldd<T, VX>(S, i) * cvt<VX>(pnew_st(i));
    return __builtin_convertvector(pnew_st(i), D);
  };
  auto unew1 = [&](long long i) -> A {
    const double p = (double)(T)((double)r[i] + beta * (double)pin[i]);
    if constexpr (SC) return (double)S[i] * p;
    return (A)p;
  };
  // registers: u of planes z-1, z, z+1 in A; p, S and kz kept as stored (T)
  D um = (A)0, uc = (A)0;
  V pc = (T)0, Sc = (T)0, kzm = (T)0;
  if (t.active) {
    const int zm0 = t.za == 0 ? nz - 1 : t.za - 1;
    long long i = (long long)zm0 * pl + row + t.x0;
    um = unew(i);
    kzm = ldv<T, VX>(kz, i);
    i = (long long)t.za * pl + row + t.x0;
    pc = pnew_own(i);
    Sc = sld(i);
    uc = cva<A, VX>(Sc) * cva<A, VX>(pc);
#pragma unroll
    for (int v = 0; v < VX; ++v) su[0][lofs + v] = uc[v];
  }
  __syncthreads();
  double dot = 0.0;
  for (int z = t.za; z < t.zb; ++z) {
    const int buf = (z - t.za) & 1;
    if (t.active) {
      const long long base = (long long)z * pl;
      const long long ic = base + row + t.x0;
      const long long in = (long long)wrapp(z, nz) * pl + row + t.x0;
      const V pn = z + 1 < t.zb ? pnew_own(in) : pnew_st(in);
      const V Sn = sld(in);
      const D un = cva<A, VX>(Sn) * cva<A, VX>(pn);
      if (z + 1 < t.zb) {
#pragma unroll
        for (int v = 0; v < VX; ++v) su[buf ^ 1][lofs + v] = un[v];
      }
      const A* sp = su[buf];
      // z and y couplings first (frees registers early)
      D l = cva<A, VX>(ldv<T, VX>(kz, ic)) * (uc - un) + cva<A, VX>(kzm) * (uc - um);
      kzm = ldv<T, VX>(kz, ic);
      {
        D uy;
        if (ym_in) {
#pragma unroll
          for (int v = 0; v < VX; ++v) uy[v] = sp[lofs - tx + v];
        } else {
          uy = unew(base + rowm + t.x0);
        }
        l += cva<A, VX>(ldv<T, VX>(ky, base + rowm + t.x0)) * (uc - uy);
        if (yp_in) {
#pragma unroll
          for (int v = 0; v < VX; ++v) uy[v] = sp[lofs + tx + v];
        } else {
          uy = unew(base + rowp + t.x0);
        }
        l += cva<A, VX>(ldv<T, VX>(ky, ic)) * (uc - uy);
      }
      {
        const A uxm = xm_in ? sp[lofs - 1] : unew1(base + row + xm);
        const A uxp = xp_in ? sp[lofs + VX] : unew1(base + row + xp);
        const D kxc = cva<A, VX>(ldv<T, VX>(kx, ic));
        const A kxm = (A)kx[base + row + xm];
#pragma unroll
        for (int v = 0; v < VX; ++v) {
          const A up = v + 1 < VX ? uc[(v + 1) % VX] : uxp;
          const A uw = v > 0 ? uc[(v + VX - 1) % VX] : uxm;
          const A kw = v > 0 ? kxc[(v + VX - 1) % VX] : kxm;
          l[v] += kxc[v] * (uc[v] - up) + kw * (uc[v] - uw);
        }
      }
      if (has_m) l += cva<A, VX>(ldv<T, VX>(mm, ic)) * uc;
      const V qv = __builtin_convertvector(cva<A, VX>(Sc) * l, V);
      stv<T, VX>(pout, ic, pc);
      stv<T, VX>(q, ic, qv);
      const vec_t<double, VX> pcd = cvt<VX>(pc), qd = cvt<VX>(qv);
#pragma unroll
      for (int v = 0; v < VX; ++v) dot += pcd[v] * qd[v];
      um = uc;
      uc = un;
      pc = pn;
      Sc = Sn;
    }
    __syncthreads();
  }
  double acc[1] = {dot};
  block_partial<1, 0u>(acc, s.partials, rc.nblk, lb);
}

// Finalize kernels: one workgroup per system, launched after the reducing
// kernel of the same name.
__global__ __launch_bounds__(kBlock) void k_fin_pcg_spmv(const SysDev* __restrict__ sds) {
  const SysDev& s = sds[blockIdx.x];
  Scalars* sc = s.sc;
  if (sc->done || sc->ru) return;  // ru: a reliable update is pending
  double acc[1];
  reduce_partials<1, 0u>(acc, s.partials, s.nblk_st);
  if (threadIdx.x == 0) {
    const double pq = acc[0];
    sc->xh_fold = sc->xh_fold == 1 ? 2 : 0;
    if (!(pq != 0.0) || !isfinite(pq)) {
      sc->status = 6;
      sc->done = 1;
    } else {
      sc->alpha = sc->rho / pq;
    }
    sc->cur ^= 1;
  }
}

// Cells [beg, end) of workgroup lb of a streaming launch of nblk workgroups:
// contiguous chunks (multiples of kBlock * VX cells), which stream from HBM
// faster than a grid-stride loop over the same workgroups.
template <int VX>
__device__ __forceinline__ void stream_range(long long n, int nblk, int lb, long long& beg,
                                             long long& end) {
  const long long unit = (long long)kBlock * VX;
  const long long chunk = (n + (long long)nblk * unit - 1) / ((long long)nblk * unit) * unit;
  beg = (long long)lb * chunk;
  end = beg + chunk < n ? beg + chunk : n;
}

// r -= alpha q ; (r,r), which decides whether a reliable update (FP64 true
// residual) follows.  x^ += alpha p is left to the CG operator of the next
// iteration, which reads p anyway (k_pcg_spmv, XH).
template <typename T, int VX>
__global__ __launch_bounds__(kBlock) void k_pcg_update(const SysDev* __restrict__ sds,
                                                       const Rec* __restrict__ recs,
                                                       const int* __restrict__ b2r) {
  using D = vec_t<double, VX>;
  const int b = blockIdx.x;
  const Rec rc = recs[b2r[b]];
  const SysDev& s = sds[rc.sys];
  Scalars* sc = s.sc;
  if (sc->done || sc->ru) return;  // ru: a reliable update is pending
  const int lb = b - rc.blk0;
  const double alpha = sc->alpha;
  const T* __restrict__ q = (const T*)s.q;
  T* __restrict__ r = (T*)s.r;
  double acc[1] = {0.0};
  long long beg, end;
  stream_range<VX>(s.n, rc.nblk, lb, beg, end);
  for (long long i = beg + (long long)threadIdx.x * VX; i < end; i += (long long)kBlock * VX) {
    const D rv = rnd<T, VX>(ldd<T, VX>(r, i) - alpha * ldd<T, VX>(q, i));
    std_<T, VX>(r, i, rv);
#pragma unroll
    for (int v = 0; v < VX; ++v) acc[0] += rv[v] * rv[v];
  }
  block_partial<1, 0u>(acc, s.partials, rc.nblk, lb);
}

__global__ __launch_bounds__(kBlock) void k_fin_pcg_update(const SysDev* __restrict__ sds) {
  const SysDev& s = sds[blockIdx.x];
  Scalars* sc = s.sc;
  if (sc->done || sc->ru) return;  // ru: a reliable update is pending
  double acc[1];
  reduce_partials<1, 0u>(acc, s.partials, s.nblk_sm);
  if (threadIdx.x == 0) {
    const double rn = acc[0];
    sc->iter += 1;
    const double est = sc->gamma * sqrt(rn);
    const bool trig = !isfinite(rn) || sqrt(rn) < sc->delta * sqrt(sc->rmax2) ||
                      est <= 0.9 * sc->tol || sc->iter > sc->maxiter;
    if (trig) {
      sc->ru = 1;
      sc->rho_pend = rn;
    } else if (!sc->pc) {
      // Jacobi-scaled CG: z = r, so (r, z) = (r, r); with a multigrid
      // preconditioner beta comes from (r, z) after the V-cycle
      sc->beta = rn / sc->rho;
      sc->rho = rn;
    }
  }
}

// Reliable update, part 1: x64 = (x64 - mean_pending) + S * x^ ; mean of
// the new x64 (nullspace removal).  x^ lacks this iteration's alpha p (the
// next iteration's CG operator would add it), so alpha p is added here and
// the operator skips it; x^ counts as zero from here on (xh_fold 1: not read
// again before the operator after next overwrites it).
template <typename T, int VX>
__global__ __launch_bounds__(kBlock) void k_ru_accum(const SysDev* __restrict__ sds,
                                                     const Rec* __restrict__ recs,
                                                     const int* __restrict__ b2r) {
  using D = vec_t<double, VX>;
  const int b = blockIdx.x;
  const Rec rc = recs[b2r[b]];
  const SysDev& s = sds[rc.sys];
  Scalars* sc = s.sc;
  // a zero start: x64 and x^ are zero already, and so is their mean
  if (sc->done || !sc->ru || sc->zero_start) return;
  const int lb = b - rc.blk0;
  const double mp = sc->mean_pending;
  const T* __restrict__ S = (const T*)s.S;
  T* __restrict__ xh = (T*)s.xh;
  double* __restrict__ x = s.x64;
  const double fa = sc->alpha;  // 0 at a start
  const bool stale = sc->xh_fold != 0;  // x^ counts as zero
  const T* __restrict__ p = (const T*)s.p[sc->cur];
  double acc[1] = {0.0};
  long long beg, end;
  stream_range<VX>(s.n, rc.nblk, lb, beg, end);
  for (long long i = beg + (long long)threadIdx.x * VX; i < end; i += (long long)kBlock * VX) {
    const D sv = S ? ldd<T, VX>(S, i) : (D)(1.0);
    D h = stale ? (D)(0.0) : ldd<T, VX>(xh, i);
    if (fa != 0.0) h += fa * ldd<T, VX>(p, i);
    const D xv = (*reinterpret_cast<const D*>(x + i) - mp) + sv * h;
    *reinterpret_cast<D*>(x + i) = xv;
#pragma unroll
    for (int v = 0; v < VX; ++v) acc[0] += xv[v];
  }
  block_partial<1, 0u>(acc, s.partials, rc.nblk, lb);
}

__global__ __launch_bounds__(kBlock) void k_fin_ru_accum(const SysDev* __restrict__ sds) {
  const SysDev& s = sds[blockIdx.x];
  Scalars* sc = s.sc;
  if (sc->done || !sc->ru) return;
  double acc[1] = {0.0};  // a zero start's sum (k_ru_accum wrote no partials)
  if (!sc->zero_start) reduce_partials<1, 0u>(acc, s.partials, s.nblk_sm);
  if (threadIdx.x == 0) {
    sc->mean_pending = sc->singular ? acc[0] / (double)s.n : 0.0;
    sc->xh_fold = 1;
  }
}

// Reliable update, part 2: r64 = b - A (x64 - mean) in FP64 with the user's
// symmetric coefficients; r^ = sign * S * r64; stop test on the true residual;
// beta from the replaced residual.  One cell per thread marching through z:
// x of planes z-1, z, z+1 and e_z of plane z-1 in registers, the x plane of
// the tile in LDS for the x and y neighbours.
template <typename T, int VX>
__global__ __launch_bounds__(kBlock) void k_ru_resid(const SysDev* __restrict__ sds,
                                                     const Rec* __restrict__ recs,
                                                     const int* __restrict__ b2r) {
  static_assert(VX == 1, "one cell per thread");
  __shared__ double su[2][kBlock];
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const Rec rc = recs[b2r[b]];
  const SysDev& s = sds[rc.sys];
  Scalars* sc = s.sc;
  if (sc->done || !sc->ru) return;
  const int lb = b - rc.blk0;
  const Tile t = tile_of<1>(rc, s, lb);
  const int nx = s.nx, ny = s.ny, nz = s.nz;
  const double mp = sc->mean_pending;
  const double sgn = s.neg ? -1.0 : 1.0;
  const gcp<double> __restrict__ x = (gcp<double>)s.x64;
  const gcp<double> __restrict__ bb = (gcp<double>)s.b64;
  const gcp<double> __restrict__ e0 = (gcp<double>)s.s4[0];
  const gcp<double> __restrict__ ex = (gcp<double>)s.s4[1];
  const gcp<double> __restrict__ ey = (gcp<double>)s.s4[2];
  const gcp<double> __restrict__ ez = (gcp<double>)s.s4[3];
  const gcp<T> __restrict__ S = (gcp<T>)s.S;
  const gmp<T> __restrict__ r = (gmp<T>)s.r;
  const int lx = rc.lx;
  const int ix = threadIdx.x % lx, iy = threadIdx.x / lx;
  const int ty = kBlock / lx;
  const bool xm_in = ix > 0;
  const bool xp_in = ix + 1 < lx && t.x0 + 1 < nx;
  const bool ym_in = iy > 0;
  const bool yp_in = iy + 1 < ty && t.y + 1 < ny;
  const long long pl = (long long)nx * ny;
  const long long row = (long long)t.y * nx;
  const long long rowm = (long long)wrapm(t.y, ny) * nx;
  const long long rowp = (long long)wrapp(t.y, ny) * nx;
  const int xw = wrapm(t.x0, nx), xe = wrapp(t.x0, nx);
  // a zero start (x64 = 0): A x is exactly zero, so b - A x = b; the same
  // cells, order and partial sums without loading x or the coefficients
  const bool zs = sc->zero_start;
  const bool e0s = s.e0_sum;  // e0 from the six neighbour coefficients
  double acc[3] = {0.0, 0.0, 0.0};
  double um = 0.0, uc = 0.0, ezm = 0.0;
  if (t.active && !zs) {
    const long long im = (long long)wrapm(t.za, nz) * pl + row + t.x0;
    um = x[im] - mp;
    ezm = ez[im];
    uc = x[(long long)t.za * pl + row + t.x0] - mp;
    su[0][threadIdx.x] = uc;
  }
  __syncthreads();
  for (int z = t.za; z < t.zb; ++z) {
    const int buf = (z - t.za) & 1;
    if (t.active && zs) {
      const long long c = (long long)z * pl + row + t.x0;
      const double r64 = bb[c];
      const T rh = (T)(sgn * (S ? (double)S[c] : 1.0) * r64);
      r[c] = rh;
      acc[0] += (double)rh * (double)rh;
      acc[1] += r64 * r64;
      acc[2] = fmax(acc[2], fabs(r64));
    } else if (t.active) {
      const long long base = (long long)z * pl;
      const long long c = base + row + t.x0;
      const double un = x[(long long)wrapp(z, nz) * pl + row + t.x0] - mp;
      if (z + 1 < t.zb) su[buf ^ 1][threadIdx.x] = un;
      const double* sp = su[buf];
      const double uxm = xm_in ? sp[threadIdx.x - 1] : x[base + row + xw] - mp;
      const double uxp = xp_in ? sp[threadIdx.x + 1] : x[base + row + xe] - mp;
      const double uym = ym_in ? sp[threadIdx.x - lx] : x[base + rowm + t.x0] - mp;
      const double uyp = yp_in ? sp[threadIdx.x + lx] : x[base + rowp + t.x0] - mp;
      const double ezc = ez[c];
      const double exw = ex[base + row + xw], exc = ex[c];
      const double eym = ey[base + rowm + t.x0], eyc = ey[c];
      const double e0c = e0s ? -(((((exw + exc) + eym) + eyc) + ezm) + ezc) : e0[c];
      double ax = e0c * uc;
      ax += exw * uxm;
      ax += exc * uxp;
      ax += eym * uym;
      ax += eyc * uyp;
      ax += ezm * um;
      ax += ezc * un;
      const double r64 = bb[c] - ax;
      const T rh = (T)(sgn * (S ? (double)S[c] : 1.0) * r64);
      r[c] = rh;
      acc[0] += (double)rh * (double)rh;
      acc[1] += r64 * r64;
      acc[2] = fmax(acc[2], fabs(r64));
      ezm = ezc;
      um = uc;
      uc = un;
    }
    if (!zs) __syncthreads();
  }
  block_partial<3, 4u>(acc, s.partials, rc.nblk, lb);
}

// k_ru_resid with 2 cells per thread (even nx): 16-byte FP64 loads of x, b
// and the coefficients; the same per-cell operation order and partials.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_ru_resid2(const SysDev* __restrict__ sds,
                                                      const Rec* __restrict__ recs,
                                                      const int* __restrict__ b2r) {
  using D2 = vec_t<double, 2>;
  __shared__ double su[2][kBlock * 2];
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const Rec rc = recs[b2r[b]];
  const SysDev& s = sds[rc.sys];
  Scalars* sc = s.sc;
  if (sc->done || !sc->ru) return;
  const int lb = b - rc.blk0;
  const Tile t = tile_of<2>(rc, s, lb);
  const int nx = s.nx, ny = s.ny, nz = s.nz;
  const double mp = sc->mean_pending;
  const double sgn = s.neg ? -1.0 : 1.0;
  const gcp<double> __restrict__ x = (gcp<double>)s.x64;
  const gcp<double> __restrict__ bb = (gcp<double>)s.b64;
  const gcp<double> __restrict__ e0 = (gcp<double>)s.s4[0];
  const gcp<double> __restrict__ ex = (gcp<double>)s.s4[1];
  const gcp<double> __restrict__ ey = (gcp<double>)s.s4[2];
  const gcp<double> __restrict__ ez = (gcp<double>)s.s4[3];
  const gcp<T> __restrict__ S = (gcp<T>)s.S;
  const gmp<T> __restrict__ r = (gmp<T>)s.r;
  const int lx = rc.lx;
  const int ix = threadIdx.x % lx, iy = threadIdx.x / lx;
  const int ty = kBlock / lx;
  const int tx = lx * 2;
  const int lofs = iy * tx + ix * 2;
  const bool xm_in = ix > 0;
  const bool xp_in = ix + 1 < lx && t.x0 + 2 < nx;
  const bool ym_in = iy > 0;
  const bool yp_in = iy + 1 < ty && t.y + 1 < ny;
  const long long pl = (long long)nx * ny;
  const long long row = (long long)t.y * nx;
  const long long rowm = (long long)wrapm(t.y, ny) * nx;
  const long long rowp = (long long)wrapp(t.y, ny) * nx;
  const int xw = wrapm(t.x0, nx), xe = t.x0 + 2 == nx ? 0 : t.x0 + 2;
  auto ld2 = [&](gcp<double> a, long long i) { return *(gcp<D2>)(a + i); };
  // a zero start (x64 = 0): b - A x = b, as in k_ru_resid
  const bool zs = sc->zero_start;
  const bool e0s = s.e0_sum;  // e0 from the six neighbour coefficients
  double acc[3] = {0.0, 0.0, 0.0};
  D2 um = 0.0, uc = 0.0, ezm = 0.0;
  if (t.active && !zs) {
    const long long im = (long long)wrapm(t.za, nz) * pl + row + t.x0;
    um = ld2(x, im) - mp;
    ezm = ld2(ez, im);
    uc = ld2(x, (long long)t.za * pl + row + t.x0) - mp;
    su[0][lofs] = uc[0];
    su[0][lofs + 1] = uc[1];
  }
  __syncthreads();
  for (int z = t.za; z < t.zb; ++z) {
    const int buf = (z - t.za) & 1;
    if (t.active && zs) {
      const long long c = (long long)z * pl + row + t.x0;
      const D2 r64 = ld2(bb, c);
      vec_t<T, 2> rh;
#pragma unroll
      for (int v = 0; v < 2; ++v) {
        rh[v] = (T)(sgn * (S ? (double)S[c + v] : 1.0) * r64[v]);
        acc[0] += (double)rh[v] * (double)rh[v];
        acc[1] += r64[v] * r64[v];
        acc[2] = fmax(acc[2], fabs(r64[v]));
      }
      *(gmp<vec_t<T, 2>>)(r + c) = rh;
    } else if (t.active) {
      const long long base = (long long)z * pl;
      const long long c = base + row + t.x0;
      const D2 un = ld2(x, (long long)wrapp(z, nz) * pl + row + t.x0) - mp;
      if (z + 1 < t.zb) {
        su[buf ^ 1][lofs] = un[0];
        su[buf ^ 1][lofs + 1] = un[1];
      }
      const double* sp = su[buf];
      const double uxm = xm_in ? sp[lofs - 1] : x[base + row + xw] - mp;
      const double uxp = xp_in ? sp[lofs + 2] : x[base + row + xe] - mp;
      D2 uym, uyp;
      if (ym_in) {
        uym[0] = sp[lofs - tx];
        uym[1] = sp[lofs - tx + 1];
      } else {
        uym = ld2(x, base + rowm + t.x0) - mp;
      }
      if (yp_in) {
        uyp[0] = sp[lofs + tx];
        uyp[1] = sp[lofs + tx + 1];
      } else {
        uyp = ld2(x, base + rowp + t.x0) - mp;
      }
      const D2 exc = ld2(ex, c), eyc = ld2(ey, c), ezc = ld2(ez, c);
      const D2 eym = ld2(ey, base + rowm + t.x0), bc = ld2(bb, c);
      const double exw = ex[base + row + xw];
      D2 e0c;
      if (e0s) {
#pragma unroll
        for (int v = 0; v < 2; ++v)
          e0c[v] = -((((((v == 0 ? exw : exc[0]) + exc[v]) + eym[v]) + eyc[v]) + ezm[v]) + ezc[v]);
      } else {
        e0c = ld2(e0, c);
      }
      D2 r64;
#pragma unroll
      for (int v = 0; v < 2; ++v) {
        const double wm = v == 0 ? uxm : uc[0], wp = v == 0 ? uc[1] : uxp;
        const double kw = v == 0 ? exw : exc[0];
        double ax = e0c[v] * uc[v];
        ax += kw * wm;
        ax += exc[v] * wp;
        ax += eym[v] * uym[v];
        ax += eyc[v] * uyp[v];
        ax += ezm[v] * um[v];
        ax += ezc[v] * un[v];
        r64[v] = bc[v] - ax;
      }
      vec_t<T, 2> rh;
#pragma unroll
      for (int v = 0; v < 2; ++v) {
        rh[v] = (T)(sgn * (S ? (double)S[c + v] : 1.0) * r64[v]);
        acc[0] += (double)rh[v] * (double)rh[v];
        acc[1] += r64[v] * r64[v];
        acc[2] = fmax(acc[2], fabs(r64[v]));
      }
      *(gmp<vec_t<T, 2>>)(r + c) = rh;
      ezm = ezc;
      um = uc;
      uc = un;
    }
    if (!zs) __syncthreads();
  }
  block_partial<3, 4u>(acc, s.partials, rc.nblk, lb);
}

__global__ __launch_bounds__(kBlock) void k_fin_ru_resid(const SysDev* __restrict__ sds) {
  const SysDev& s = sds[blockIdx.x];
  Scalars* sc = s.sc;
  if (sc->done || !sc->ru) return;
  double acc[3];
  reduce_partials<3, 4u>(acc, s.partials, s.nblk_ru);
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
    } else if (sc->iter >= sc->miniter && (conv || sc->iter > sc->maxiter)) {
      // the reference's rule (src/linear/linear.ipp:109-110): at most
      // maxiter + 1 iterations, never fewer than miniter
      sc->converged = conv ? 1 : 0;
      sc->done = 1;
    } else {
      if (!sc->pc) {
        sc->beta = sc->first ? 0.0 : rh2 / sc->rho;
        sc->rho = rh2;
      }
      sc->rmax2 = rh2;
      sc->gamma = rh2 > 0 ? metric / sqrt(rh2) : 1.0;
      if (!sc->pc) sc->first = 0;
    }
    sc->ru = 0;
    sc->zero_start = 0;
  }
}

// ===========================================================================
// Multigrid preconditioner z = M^-1 r: one symmetric V-cycle over the levels
// of each system.  Level operators (L u)_c = sum_f k_f (u_c - u_nb) + m_c u_c
// with the coarse conductances from coarsening (k_mg_coarsen); smoothing by
// two damped Jacobi sweeps before and two after the coarse correction, with
// the weights kOmegaA, kOmegaB of the degree-2 Chebyshev polynomial for the
// spectrum interval [kChebLo, kChebHi] of D^-1 L (w = 1 / (c -+ d cos(pi/4)),
// c = (lo + hi) / 2, d = (hi - lo) / 2): pre-smoothing A then B (A from a
// zero guess), post-smoothing B then A; both apply the same polynomial of
// D^-1 L, so the cycle stays symmetric.  The coarsest level: `sweeps` sweeps
// of weight kOmega.  Restriction = sum over the 2x2x2 children, prolongation
// = injection (its transpose) scaled by kProlong (an over-correction of the
// piecewise-constant interpolation; a scaled symmetric term), so the V-cycle
// is a symmetric linear operator.  Arithmetic in the engine's element type T.
// ===========================================================================
constexpr double kOmega = 0.9;
constexpr double kChebLo = 0.5, kChebHi = 1.8;
constexpr double kOmegaA = 0.6212648748965077;  // 1 / (c + d cos(pi/4))
constexpr double kOmegaB = 1.4484764074432;     // 1 / (c - d cos(pi/4))
constexpr double kChebC = 0.5 * (kChebHi + kChebLo), kChebD = 0.5 * (kChebHi - kChebLo);
constexpr double kCosQ = 0.70710678118654752;  // cos(pi/4)
static_assert(kOmegaA * (kChebC + kChebD * kCosQ) - 1.0 < 1e-12 &&
                  1.0 - kOmegaA * (kChebC + kChebD * kCosQ) < 1e-12,
              "kOmegaA is the first Chebyshev weight of [kChebLo, kChebHi]");
static_assert(kOmegaB * (kChebC - kChebD * kCosQ) - 1.0 < 1e-12 &&
                  1.0 - kOmegaB * (kChebC - kChebD * kCosQ) < 1e-12,
              "kOmegaB is the second Chebyshev weight of [kChebLo, kChebHi]");
constexpr double kProlong = 1.1;

template <typename T>
struct LevelView {
  int nx, ny, nz;
  long long pl;
  MgCoefs<T> co;
  const T* m;
  __device__ explicit LevelView(const MgLevel& L)
      : nx(L.nx), ny(L.ny), nz(L.nz), pl((long long)L.nx * L.ny), co(L), m((const T*)L.m) {}
  __device__ T kx(long long i) const { return co.comp1(0, i); }
  __device__ T ky(long long i) const { return co.comp1(1, i); }
  __device__ T kz(long long i) const { return co.comp1(2, i); }
  __device__ T dinv(long long i) const { return co.comp1(3, i); }
  __device__ long long at(int x, int y, int z) const { return (long long)z * pl + (long long)y * nx + x; }
  // (L u)_c where U(i) returns u at flat index i
  template <class U>
  __device__ T lap(int x, int y, int z, U u) const {
    const long long c = at(x, y, z);
    const long long xm = at(wrapm(x, nx), y, z), xp = at(wrapp(x, nx), y, z);
    const long long ym = at(x, wrapm(y, ny), z), yp = at(x, wrapp(y, ny), z);
    const long long zm = at(x, y, wrapm(z, nz)), zp = at(x, y, wrapp(z, nz));
    const T uc = u(c);
    T s = kx(c) * (uc - u(xp)) + kx(xm) * (uc - u(xm)) + ky(c) * (uc - u(yp)) +
          ky(ym) * (uc - u(ym)) + kz(c) * (uc - u(zp)) + kz(zm) * (uc - u(zm));
    if (m) s += m[c] * uc;
    return s;
  }
};

// Explicit address spaces for the bottom kernel: its level data sit in LDS
// (or, for levels too large for it, in global memory) behind generic
// pointers, which would otherwise compile to flat loads with 64-bit
// addresses.
template <typename T>
using lds_cp = const __attribute__((address_space(3))) T*;
template <typename T>
using lds_p = __attribute__((address_space(3))) T*;
template <typename T>
using glb_cp = const __attribute__((address_space(1))) T*;
template <typename T>
using glb_p = __attribute__((address_space(1))) T*;

// A level of the bottom kernel with 32-bit indices: the operator and
// operation order of LevelView, periodic neighbours from the flat index.
// CL: coefficients, mass and right-hand side in LDS (else global memory).
template <typename T, bool CL>
struct LdsView {
  template <typename U>
  using CP = std::conditional_t<CL, lds_cp<U>, glb_cp<U>>;
  template <typename U>
  using P = std::conditional_t<CL, lds_p<U>, glb_p<U>>;
  using E = std::conditional_t<sizeof(T) == 4, unsigned short, T>;
  int nx, ny, nz, pl, n;
  CP<E> a[4];
  CP<T> m;
  bool hasm;
  P<T> r;
  __device__ explicit LdsView(const MgLevel& L)
      : nx(L.nx), ny(L.ny), nz(L.nz), pl(L.nx * L.ny), n((int)L.n), hasm(L.m != nullptr) {
    if constexpr (sizeof(T) == 4) {
      a[0] = a[1] = a[2] = a[3] = (CP<E>)(const E*)L.kc;
    } else {
      a[0] = (CP<E>)(const E*)L.kx;
      a[1] = (CP<E>)(const E*)L.ky;
      a[2] = (CP<E>)(const E*)L.kz;
      a[3] = (CP<E>)(const E*)L.dinv;
    }
    m = hasm ? (CP<T>)(const T*)L.m : CP<T>{};
    r = (P<T>)(T*)L.r;
  }
  __device__ T k(int c, int i) const {
    if constexpr (sizeof(T) == 4) {
      return __uint_as_float((unsigned)a[0][((i >> 2) << 4) + 4 * c + (i & 3)] << 16);
    } else {
      return a[c][i];
    }
  }
  __device__ T dinv(int i) const { return k(3, i); }
  // (L u)_c of cell c = (x, y, z)
  __device__ T lap(int c, int x, int y, int z, lds_cp<T> u) const {
    const int xm = x == 0 ? c + nx - 1 : c - 1, xp = x == nx - 1 ? c - nx + 1 : c + 1;
    const int ym = y == 0 ? c + pl - nx : c - nx, yp = y == ny - 1 ? c - pl + nx : c + nx;
    const int zm = z == 0 ? c + n - pl : c - pl, zp = z == nz - 1 ? c - n + pl : c + pl;
    const T uc = u[c];
    T s = k(0, c) * (uc - u[xp]) + k(0, xm) * (uc - u[xm]) + k(1, c) * (uc - u[yp]) +
          k(1, ym) * (uc - u[ym]) + k(2, c) * (uc - u[zp]) + k(2, zm) * (uc - u[zm]);
    if (hasm) s += m[c] * uc;
    return s;
  }
};

// f(i, x, y, z) for the cells i = threadIdx.x + j * blockDim.x < n of an
// (nx, ny, .) grid; the coordinates advance by the block's stride without
// divisions.
template <class F>
__device__ __forceinline__ void for_cells(int n, int nx, int ny, F f) {
  const int nt = blockDim.x, pl = nx * ny;
  int i = threadIdx.x;
  if (i >= n) return;
  int z = i / pl, y = (i - z * pl) / nx, x = i - z * pl - y * nx;
  const int sz = nt / pl, sy = (nt - sz * pl) / nx, sx = nt - sz * pl - sy * nx;
  for (; i < n; i += nt) {
    f(i, x, y, z);
    x += sx;
    y += sy;
    z += sz;
    if (x >= nx) {
      x -= nx;
      ++y;
    }
    if (y >= ny) {
      y -= ny;
      ++z;
    }
  }
}

// Cell of a one-cell-per-thread stencil tile of level grid (nx, ny, nz).
struct Cell {
  int x, y, za, zb;
  bool active;
};

__device__ __forceinline__ Cell cell_of(const Rec& rc, int lb, int nx, int ny, int nz) {
  Cell c;
  const int tx = lb % rc.tiles_x;
  const int t2 = lb / rc.tiles_x;
  const int ty = t2 % rc.tiles_y;
  const int tz = t2 / rc.tiles_y;
  const int rows = kBlock / rc.lx;
  const int ix = threadIdx.x % rc.lx, iy = threadIdx.x / rc.lx;
  c.x = tx * rc.lx + ix;
  c.y = ty * rows + iy;
  c.za = tz * rc.bz;
  c.zb = min(c.za + rc.bz, nz);
  c.active = iy < rows && c.x < nx && c.y < ny;
  return c;
}

// Smoothing sweep (the bandwidth-bound levels): each thread owns VX
// consecutive cells of a row and marches through z with u of planes z-1, z,
// z+1 in registers; the tile's u plane is shared through LDS for the x and y
// neighbours (edge threads recompute out-of-tile neighbours from memory).
//   MODE 0: u = z_in
//   MODE 1: u = w D^-1 r (the zero-guess first sweep, fused)
//   MODE 2: u = z_in + P z_c (coarse correction, fused)
// and z_out = u + w D^-1 (r - L u); DOT: partials of (r, z_out).
//   addzs (MODE 0, the last sweep of a second coarse correction): z_out also
//           adds zs, the first correction, so that the prolongation above
//           reads one coarse array

//   MODE 4: u = z_in, no z_out: r = r - L u in place and u saved to zs (the
//           start of the second coarse correction of the W level above; no
//           copy on fused and tile levels, whose zs is z[1] itself)
//   MODE 3: u = z_in, no z_out: the residual r - L u restricted to level
//           l+1 (sum over the 2x2x2 children); x pairs are summed in the
//           thread (VX >= 2), plane pairs over the z march (even plane count
//           per tile) and row pairs through LDS.
// L0: the launch on level 0 (a separate instance, so that profiles keep the
// bandwidth-bound fine-grid sweeps apart from the coarse levels).
template <typename T, int VX, int MODE, bool DOT, bool L0>
__global__ __launch_bounds__(kBlock) void k_mg_sweep(const SysDev* __restrict__ sds,
                                                     const Rec* __restrict__ recs,
                                                     const int* __restrict__ b2r, int l_arg,
                                                     int zin, int zout, int wl, int so,
                                                     int addzs) {
  const int l = L0 ? 0 : l_arg;
  using V = vec_t<T, VX>;
  static_assert(MODE != 3 || VX >= 2, "the fused restriction pairs cells in the thread");
  constexpr int NP = VX >= 2 ? VX / 2 : 1;  // coarse cells per thread (MODE 3)
  __shared__ T su[2][kBlock * VX];
  __shared__ T sacc[MODE == 3 ? kBlock * NP : 1];
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const Rec rc = recs[b2r[b]];
  const SysDev& s = sds[rc.sys];
  if (s.sc->done) return;
  // a launch of the second coarse correction below W level wl: only the
  // systems with more than wl W levels take part
  if (wl >= 0 && wl >= s.mg_w) return;
  // so: the start of newly prepared systems (first == 1); running ones skip
  if (so && !s.sc->first) return;
  const MgLevel& Lv = s.lev[l];
  // this system's level runs fused kernels (k_mg_pre, k_mg_post) instead
  if ((MODE == 0 || MODE == 1 || MODE == 2 || MODE == 3) && Lv.post) return;
  // zin / zout < 0: the level's pre-smoothed iterate (z[0] if fused, else z[1])
  if (zin < 0) zin = Lv.post ? 0 : 1;
  if (zout < 0) zout = Lv.post ? 0 : 1;
  const int nx = Lv.nx, ny = Lv.ny, nz = Lv.nz;
  const int lb = b - rc.blk0;
  // tile of the level grid
  const int tix = lb % rc.tiles_x;
  const int t2 = lb / rc.tiles_x;
  const int tiy = t2 % rc.tiles_y;
  const int tiz = t2 / rc.tiles_y;
  const int lx = rc.lx;
  const int ty = kBlock / lx;
  const int ix = threadIdx.x % lx, iy = threadIdx.x / lx;
  const int x0 = (tix * lx + ix) * VX;
  const int y = tiy * ty + iy;
  const int za = tiz * rc.bz, zb = min(za + rc.bz, nz);
  const bool active = iy < ty && x0 < nx && y < ny;
  const int tx = lx * VX;
  const int lofs = iy * tx + ix * VX;
  const bool xm_in = ix > 0;
  const bool xp_in = ix + 1 < lx && x0 + VX < nx;
  const bool ym_in = iy > 0;
  const bool yp_in = iy + 1 < ty && y + 1 < ny;
  const MgCoefs<T> co(Lv);
  const gcp<T> __restrict__ mm = (gcp<T>)Lv.m;
  const gcp<T> __restrict__ r = (gcp<T>)Lv.r;
  const gcp<T> __restrict__ zi = (gcp<T>)Lv.z[zin];
  const gmp<T> __restrict__ zo = (gmp<T>)Lv.z[zout];
  gcp<T> __restrict__ zc = nullptr;
  gmp<T> __restrict__ rco = nullptr;
  int cnx = 0, cny = 0;
  if constexpr (MODE == 2 || MODE == 3) {
    const MgLevel& Lc = s.lev[l + 1];
    if constexpr (MODE == 2) {
      zc = (gcp<T>)Lc.z[l + 1 == s.nlev - 1 ? 0 : 1];
    } else {
      rco = (gmp<T>)Lc.r;
    }
    cnx = Lc.nx;
    cny = Lc.ny;
  }
  gmp<T> __restrict__ rw = nullptr;   // MODE 4: residual written in place
  gmp<T> __restrict__ zsv = nullptr;  // MODE 4: the correction saved
  gcp<T> __restrict__ zsa = nullptr;  // MODE 0: the first correction added
  if constexpr (MODE == 4) {
    rw = (gmp<T>)Lv.r;
    // no copy where zs is z_in itself (fused and tile levels)
    if (Lv.zs != Lv.z[zin]) zsv = (gmp<T>)Lv.zs;
  }
  if constexpr (MODE == 0) {
    if (addzs) zsa = (gcp<T>)Lv.zs;
  }
  // MODE 1: zero guess (weight A) + sweep B; MODE 2: sweep B; MODE 0: sweep A
  const T wa = (T)kOmegaA;
  const T w = MODE == 0 ? (T)kOmegaA : (T)kOmegaB;
  const long long pl = (long long)nx * ny;
  const long long row = (long long)y * nx;
  const int ym = wrapm(y, ny), yp = wrapp(y, ny);
  const int xm = x0 == 0 ? nx - 1 : x0 - 1;
  const int xp = x0 + VX == nx ? 0 : x0 + VX;
  // u at (x, yy, zz) for VX cells / for one cell
  auto uv = [&](int xx, int yy, int zz) -> V {
    const long long i = (long long)zz * pl + (long long)yy * nx + xx;
    if constexpr (MODE == 1) {
      return wa * co.template comp<VX>(3, i) * ldv<T, VX>(r, i);
    } else if constexpr (MODE == 2) {
      V u = ldv<T, VX>(zi, i);
      const long long cb = ((long long)(zz >> 1) * cny + (yy >> 1)) * cnx + (xx >> 1);
      if constexpr (VX >= 2) {
        // xx is a multiple of VX: the VX / 2 coarse cells are contiguous
        const vec_t<T, VX / 2> c = ldv<T, VX / 2>(zc, cb);
#pragma unroll
        for (int v = 0; v < VX; ++v) u[v] += (T)kProlong * c[v / 2];
      } else {
        u[0] += (T)kProlong * zc[cb];
      }
      return u;
    } else {
      return ldv<T, VX>(zi, i);
    }
  };
  auto u1 = [&](int xx, int yy, int zz) -> T {
    const long long i = (long long)zz * pl + (long long)yy * nx + xx;
    if constexpr (MODE == 1) {
      return wa * co.comp1g(3, i) * ld1(r, i);
    } else if constexpr (MODE == 2) {
      const long long ci = ((long long)(zz >> 1) * cny + (yy >> 1)) * cnx + (xx >> 1);
      return zi[i] + (T)kProlong * zc[ci];
    } else {
      return zi[i];
    }
  };
  V um = (T)0, uc = (T)0, kzm = (T)0;
  // MODE 1: r of the current plane, loaded one plane ahead with u
  V rcur = (T)0;
  if (active) {
    um = uv(x0, y, wrapm(za, nz));
    kzm = co.template comp<VX>(2, (long long)wrapm(za, nz) * pl + row + x0);
    if constexpr (MODE == 1) {
      const long long i0 = (long long)za * pl + row + x0;
      rcur = ldv<T, VX>(r, i0);
      uc = wa * co.template comp<VX>(3, i0) * rcur;
    } else {
      uc = uv(x0, y, za);
    }
#pragma unroll
    for (int v = 0; v < VX; ++v) su[0][lofs + v] = uc[v];
  }
  __syncthreads();
  double dot = 0.0;
  T racc[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) racc[j] = (T)0;
  for (int z = za; z < zb; ++z) {
    const int buf = (z - za) & 1;
    if (active) {
      const long long base = (long long)z * pl;
      const long long ic = base + row + x0;
      V un, rnext = (T)0;
      if constexpr (MODE == 1) {
        const long long in = (long long)wrapp(z, nz) * pl + row + x0;
        rnext = ldv<T, VX>(r, in);
        un = wa * co.template comp<VX>(3, in) * rnext;
      } else {
        un = uv(x0, y, wrapp(z, nz));
      }
      if (z + 1 < zb) {
#pragma unroll
        for (int v = 0; v < VX; ++v) su[buf ^ 1][lofs + v] = un[v];
      }
      const T* sp = su[buf];
      V kxc, kyc, kzc, dc;
      co.template own<VX>(ic, kxc, kyc, kzc, dc);
      V lap = kzc * (uc - un) + kzm * (uc - um);
      kzm = kzc;
      {
        V uy;
        if (ym_in) {
#pragma unroll
          for (int v = 0; v < VX; ++v) uy[v] = sp[lofs - tx + v];
        } else {
          uy = uv(x0, ym, z);
        }
        lap += co.template comp<VX>(1, base + (long long)ym * nx + x0) * (uc - uy);
        if (yp_in) {
#pragma unroll
          for (int v = 0; v < VX; ++v) uy[v] = sp[lofs + tx + v];
        } else {
          uy = uv(x0, yp, z);
        }
        lap += kyc * (uc - uy);
      }
      {
        const T uxm = xm_in ? sp[lofs - 1] : u1(xm, y, z);
        const T uxp = xp_in ? sp[lofs + VX] : u1(xp, y, z);
        const T kxm = co.comp1g(0, base + row + xm);
#pragma unroll
        for (int v = 0; v < VX; ++v) {
          const T up = v + 1 < VX ? uc[(v + 1) % VX] : uxp;
          const T uw = v > 0 ? uc[(v + VX - 1) % VX] : uxm;
          const T kw = v > 0 ? kxc[(v + VX - 1) % VX] : kxm;
          lap[v] += kxc[v] * (uc[v] - up) + kw * (uc[v] - uw);
        }
      }
      if (mm) lap += ldv<T, VX>(mm, ic) * uc;
      V rv;
      if constexpr (MODE == 1) rv = rcur;
      else rv = ldv<T, VX>(r, ic);
      if constexpr (MODE == 3) {
        const V res = rv - lap;
#pragma unroll
        for (int j = 0; j < NP; ++j) racc[j] += res[2 * j] + res[2 * j + 1];
      } else if constexpr (MODE == 4) {
        if (zsv) stv<T, VX>(zsv, ic, uc);
        stv<T, VX>(rw, ic, rv - lap);
      } else {
        V out = uc + w * dc * (rv - lap);
        if constexpr (MODE == 0) {
          if (zsa) out += ldv<T, VX>(zsa, ic);
        }
        stv<T, VX>(zo, ic, out);
        if constexpr (DOT) {
#pragma unroll
          for (int v = 0; v < VX; ++v) dot += (double)rv[v] * (double)out[v];
        }
      }
      um = uc;
      uc = un;
      if constexpr (MODE == 1) rcur = rnext;
    }
    if constexpr (MODE == 3) {
      // second plane of a pair (za is even): odd rows hand their sums to
      // the even row above through LDS; the even row writes the coarse cells
      if ((z - za) & 1) {
        if (iy & 1) {
#pragma unroll
          for (int j = 0; j < NP; ++j) sacc[threadIdx.x * NP + j] = racc[j];
        }
        __syncthreads();
        if (active && !(iy & 1)) {
          const long long co =
              ((long long)(z >> 1) * cny + (y >> 1)) * cnx + (x0 >> 1);
#pragma unroll
          for (int j = 0; j < NP; ++j) rco[co + j] = racc[j] + sacc[(threadIdx.x + lx) * NP + j];
        }
#pragma unroll
        for (int j = 0; j < NP; ++j) racc[j] = (T)0;
      }
    }
    __syncthreads();
  }
  if constexpr (DOT) {
    double acc[1] = {dot};
    block_partial<1, 0u>(acc, s.partials, rc.nblk, lb);
  }
}

// Pre- or post-smoothing of a small level (MgLevel::post == 2, FP32 engines,
// nx % 4 == 0) in one launch: a workgroup owns an 8x8x8 tile of output cells
// and recomputes what its two sweeps need around it, so the level costs one
// launch and one round of global loads instead of two z-marching launches
// with a barrier and a load round per plane.
//   PRE:  u = wA D^-1 r (12^3 box); v = u + wB D^-1 (r - L u) (10^3 box)
//         -> z[0] on the tile; r_{l+1}[C] = sum over the 2x2x2 children of
//         (r - L v)
//   POST: u = z[0] + kProlong P z_c (12^3 box); v = u + wB D^-1 (r - L u)
//         (10^3 box); z[1] = v + wA D^-1 (r - L v) (+ zs: the last sweep of
//         a second coarse correction) on the tile
// The box is staged in LDS by rows of 16 cells (x0-4 .. x0+11, four aligned
// groups of 4 cells: r as float4, the packed bfloat16 coefficients as two
// 16-byte halves, u formed as the row is loaded), y and z from -2 to +9
// around the tile.  Rows and groups wrap periodically (a level edge shorter
// than the box repeats cells); tile cells past the level's edge are neither
// stored nor restricted (edges are even, tiles start at multiples of 8).
template <bool PRE>
__global__ __launch_bounds__(kTileBlock) void k_mg_tile(const SysDev* __restrict__ sds,
                                                        const Rec* __restrict__ recs,
                                                        const int* __restrict__ b2r, int l,
                                                        int wl, int so, int addzs) {
  constexpr int B = kTile + 4;            // box rows per edge (y, z)
  constexpr int W = kTile + 8;            // cells per staged row (x)
  constexpr int NR = B * B;               // staged rows
  constexpr int NG = W / 4;               // 4-cell groups per row
  constexpr int NT = kTile * kTile * kTile;
  __shared__ __align__(16) float sR[NR * W];
  __shared__ __align__(16) float sU[NR * W];
  __shared__ __align__(16) float sV[NR * W];
  __shared__ __align__(16) unsigned short sK[NR * NG * 16];  // kx[4] ky[4] kz[4] dinv[4] per group
  __shared__ float sres[PRE ? NT : 1];
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const Rec rc = recs[b2r[b]];
  const SysDev& s = sds[rc.sys];
  if (s.sc->done) return;
  if (wl >= 0 && wl >= s.mg_w) return;
  if (so && !s.sc->first) return;
  const MgLevel& Lv = s.lev[l];
  const MgLevel& Lc = s.lev[l + 1];
  const int nx = Lv.nx, ny = Lv.ny, nz = Lv.nz;
  const int lb = b - rc.blk0;
  const int x0 = (lb % rc.tiles_x) * kTile;
  const int t2 = lb / rc.tiles_x;
  const int y0 = (t2 % rc.tiles_y) * kTile;
  const int z0 = (t2 / rc.tiles_y) * kTile;
  const int cnx = Lc.nx, cny = Lc.ny;
  const gcp<float> r = (gcp<float>)Lv.r;
  const gcp<float> mm = (gcp<float>)Lv.m;
  const gcp<unsigned short> kc = (gcp<unsigned short>)Lv.kc;
  auto wrap = [](int v, int n) {
    v %= n;
    return v < 0 ? v + n : v;
  };
  // staging: one (row, group) unit per thread and pass
  for (int u = threadIdx.x; u < NR * NG; u += kTileBlock) {
    const int ri = u / NG, g = u % NG;
    const int yi = ri % B, zi = ri / B;
    const int y = wrap(y0 - 2 + yi, ny), z = wrap(z0 - 2 + zi, nz);
    const int gx = wrap(x0 - 4 + 4 * g, nx);  // nx % 4 == 0: a whole group
    const int i = (z * ny + y) * nx + gx;
    const vec_t<float, 4> rv = *(gcp<vec_t<float, 4>>)(r + i);
    const gcp<vec_t<unsigned, 4>> kg = (gcp<vec_t<unsigned, 4>>)(kc + ((long long)(i >> 2) << 4));
    const vec_t<unsigned, 4> k0 = kg[0], k1 = kg[1];
    vec_t<float, 4> uv;
    if constexpr (PRE) {
      // u = wA * dinv * r, dinv from the group's upper half
      const float d[4] = {__uint_as_float(k1.z << 16), __uint_as_float(k1.z & 0xffff0000u),
                          __uint_as_float(k1.w << 16), __uint_as_float(k1.w & 0xffff0000u)};
#pragma unroll
      for (int v = 0; v < 4; ++v) uv[v] = (float)kOmegaA * d[v] * rv[v];
    } else {
      const vec_t<float, 4> zv = *(gcp<vec_t<float, 4>>)((gcp<float>)Lv.z[0] + i);
      const vec_t<float, 2> cv = *(gcp<vec_t<float, 2>>)((gcp<float>)Lc.z[l + 1 == s.nlev - 1 ? 0 : 1] +
                                                          ((z >> 1) * cny + (y >> 1)) * cnx + (gx >> 1));
#pragma unroll
      for (int v = 0; v < 4; ++v) uv[v] = zv[v] + (float)kProlong * cv[v / 2];
    }
    const int o = ri * W + 4 * g;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      sR[o + v] = rv[v];
      sU[o + v] = uv[v];
    }
    vec_t<unsigned, 4>* kd = (vec_t<unsigned, 4>*)(sK + (ri * NG + g) * 16);
    kd[0] = k0;
    kd[1] = k1;
  }
  // coefficient c of the cell at row ri, column xi
  auto kf = [&](int ri, int xi, int c) {
    return __uint_as_float((unsigned)sK[(ri * NG + (xi >> 2)) * 16 + 4 * c + (xi & 3)] << 16);
  };
  // mass term of the cell (global; absent from the periodic pressure system)
  auto mass = [&](int ri, int xi) {
    const int y = wrap(y0 - 2 + ri % B, ny), z = wrap(z0 - 2 + ri / B, nz);
    return mm[(z * ny + y) * nx + wrap(x0 - 4 + xi, nx)];
  };
  // sum_f k_f (w_c - w_nb) + m w_c at (ri, xi), w staged like sR: the
  // operation order of the z-marching sweeps
  auto lap = [&](const float* w, int ri, int xi) {
    const int j = ri * W + xi;
    const float wc = w[j];
    float l7 = kf(ri, xi, 2) * (wc - w[j + B * W]) + kf(ri - B, xi, 2) * (wc - w[j - B * W]);
    l7 += kf(ri - 1, xi, 1) * (wc - w[j - W]);
    l7 += kf(ri, xi, 1) * (wc - w[j + W]);
    l7 += kf(ri, xi, 0) * (wc - w[j + 1]) + kf(ri, xi - 1, 0) * (wc - w[j - 1]);
    if (mm) l7 += mass(ri, xi) * wc;
    return l7;
  };
  __syncthreads();
  // second sweep (weight B) on the box rows 1..10, columns 3..12
  constexpr int BV = kTile + 2;
  for (int c = threadIdx.x; c < BV * BV * BV; c += kTileBlock) {
    const int xi = 3 + c % BV, ri = (1 + (c / BV) % BV) + (1 + c / (BV * BV)) * B;
    const int j = ri * W + xi;
    sV[j] = sU[j] + (float)kOmegaB * kf(ri, xi, 3) * (sR[j] - lap(sU, ri, xi));
  }
  __syncthreads();
  gmp<float> zout = (gmp<float>)(PRE ? Lv.z[0] : Lv.z[1]);
  for (int c = threadIdx.x; c < NT; c += kTileBlock) {
    const int a = c % kTile, bb = (c / kTile) % kTile, cz = c / (kTile * kTile);
    const int xi = 4 + a, ri = (2 + bb) + (2 + cz) * B;
    const int j = ri * W + xi;
    const bool in = x0 + a < nx && y0 + bb < ny && z0 + cz < nz;
    const int g = ((z0 + cz) * ny + y0 + bb) * nx + x0 + a;
    if constexpr (PRE) {
      sres[c] = sR[j] - lap(sV, ri, xi);
      if (in) zout[g] = sV[j];
    } else {
      float out = sV[j] + (float)kOmegaA * kf(ri, xi, 3) * (sR[j] - lap(sV, ri, xi));
      if (addzs && in) out += ((gcp<float>)Lv.zs)[g];
      if (in) zout[g] = out;
    }
  }
  if constexpr (PRE) {
    __syncthreads();
    // 64 coarse cells: x pairs, then plane pairs, then row pairs
    constexpr int H = kTile / 2;
    if (threadIdx.x < H * H * H) {
      const int a = threadIdx.x % H, bb = (threadIdx.x / H) % H, cz = threadIdx.x / (H * H);
      auto xp = [&](int dy, int dz) {
        const int i = ((2 * cz + dz) * kTile + 2 * bb + dy) * kTile + 2 * a;
        return sres[i] + sres[i + 1];
      };
      const float v = (xp(0, 0) + xp(0, 1)) + (xp(1, 0) + xp(1, 1));
      const int cx = x0 / 2 + a, cy = y0 / 2 + bb, ccz = z0 / 2 + cz;
      if (2 * cx < nx && 2 * cy < ny && 2 * ccz < nz)
        ((gmp<float>)Lc.r)[(ccz * cny + cy) * cnx + cx] = v;
    }
  }
}

// Residual of level l restricted to level l+1: r_{l+1}[C] = sum over the
// 2x2x2 children c of (r - L z)_c.  One thread per coarse cell.
template <typename T, bool L0>
__global__ __launch_bounds__(kBlock) void k_mg_restrict(const SysDev* __restrict__ sds,
                                                        const Rec* __restrict__ recs,
                                                        const int* __restrict__ b2r, int l_arg,
                                                        int zin, int wl, int so) {
  const int l = L0 ? 0 : l_arg;
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const Rec rc = recs[b2r[b]];
  const SysDev& s = sds[rc.sys];
  if (s.sc->done) return;
  if (wl >= 0 && wl >= s.mg_w) return;
  if (so && !s.sc->first) return;
  const MgLevel& Lf = s.lev[l];
  if (Lf.post) return;  // restricted by k_mg_pre
  const MgLevel& Lc = s.lev[l + 1];
  const LevelView<T> F(Lf);
  const int lb = b - rc.blk0;
  const Cell c = cell_of(rc, lb, Lc.nx, Lc.ny, Lc.nz);
  if (!c.active) return;
  const T* __restrict__ r = (const T*)Lf.r;
  const T* __restrict__ zi = (const T*)Lf.z[zin >= 0 ? zin : (Lf.post ? 0 : 1)];
  T* __restrict__ rcoarse = (T*)Lc.r;
  auto u = [&](long long i) -> T { return zi[i]; };
  for (int z = c.za; z < c.zb; ++z) {
    T sum = 0;
    for (int dz = 0; dz < 2; ++dz)
      for (int dy = 0; dy < 2; ++dy)
        for (int dx = 0; dx < 2; ++dx) {
          const int fx = 2 * c.x + dx, fy = 2 * c.y + dy, fz = 2 * z + dz;
          sum += r[F.at(fx, fy, fz)] - F.lap(fx, fy, fz, u);
        }
    rcoarse[((long long)z * Lc.ny + c.y) * Lc.nx + c.x] = sum;
  }
}

// Both post-smoothing sweeps of a level in one z march:
//   u = z[0] + P z_c;  v = u + wB D^-1 (r - L u);  z[1] = v + wA D^-1 (r - L v)
// (+ zs with addzs, the first correction of a W level's second correction;
// DOT: partials of (r, z[1])).  The pre-smoothing of such a level (post = 1)
// writes z[0], so z[1] ends with the correction, as after the two separate
// sweeps.
// A workgroup covers whole rows (rc.lx = nx / VX threads per row) and rc.bz
// planes.  Its rc.ty rows of threads are the output rows 1 .. ty-2 and one
// halo row on each side: v is computed on all ty rows (the halo rows' v is
// recomputed rather than exchanged between workgroups), z[1] on the output
// rows one plane behind v.  Step k of the march forms v on plane k and then
// z[1] on plane k-1: the plane neighbours of v on plane k-1 are in LDS since
// the previous step, the cell's own v on planes k-2 .. k and its coefficients
// and r on plane k-1 (loaded by the previous step's v sweep) are in
// registers, so one barrier per step suffices (u and v planes double-buffered
// in LDS) and the second sweep re-reads from memory only the coefficients of
// the neighbouring rows and columns, a plane after their first read (L2
// hits).  The u rows just outside the tile (for the halo rows' v) and the u
// planes before and after the tile come from memory.
template <typename T, int VX, bool DOT, bool L0>
__global__ __launch_bounds__(kPostBlock) void k_mg_post(const SysDev* __restrict__ sds,
                                                        const Rec* __restrict__ recs,
                                                        const int* __restrict__ b2r, int l_arg,
                                                        int wl, int so, int addzs) {
  const int l = L0 ? 0 : l_arg;
  using V = vec_t<T, VX>;
  __shared__ T su[2][kPostBlock * VX];
  __shared__ T sv[2][kPostBlock * VX];
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  // launch tables (recs, b2r, sds) and level tables (s.lev) are device
  // arrays written by the host between launches only
  const Rec rc = cload(gtab(recs + cload(gtab(b2r + b))));
  const SysDev s = cload(gtab(sds + rc.sys));
  if (s.sc->done) return;
  if (wl >= 0 && wl >= s.mg_w) return;
  if (so && !s.sc->first) return;
  const MgLevel Lv = cload(gtab(s.lev + l));
  const MgLevel Lc = cload(gtab(s.lev + l + 1));
  const int nx = Lv.nx, ny = Lv.ny, nz = Lv.nz;
  const int lb = b - rc.blk0;
  const int lx = rc.lx, ty = rc.ty;  // lx * VX == nx, ty <= kPostBlock / lx
  const int tiy = lb % rc.tiles_y, tiz = lb / rc.tiles_y;
  const int ix = threadIdx.x % lx, iy = threadIdx.x / lx;
  const bool active = iy < ty;
  // grid row of this thread row: output rows y0 .. y0+ty-3, halo rows around
  // (yr in [-1, 2 ny - 1] because ty <= ny + 2)
  const int y0 = tiy * (ty - 2);
  const int yr = y0 - 1 + iy;
  const int y = yr < 0 ? yr + ny : (yr >= ny ? yr - ny : yr);
  const bool outrow = active && iy >= 1 && iy <= ty - 2 && yr < ny;
  const int za = tiz * rc.bz, zb = min(za + rc.bz, nz);
  const int x0 = ix * VX;
  const int lofs = iy * nx + x0;
  const int lxm = iy * nx + (ix == 0 ? nx - 1 : x0 - 1);
  const int lxp = iy * nx + (ix == lx - 1 ? 0 : x0 + VX);
  const int xm = x0 == 0 ? nx - 1 : x0 - 1;
  const int ym = wrapm(y, ny), yp = wrapp(y, ny);
  const MgCoefs<T> co(Lv);
  const T* __restrict__ mm = (const T*)Lv.m;
  const T* __restrict__ r = (const T*)Lv.r;
  const T* __restrict__ zi = (const T*)Lv.z[0];
  T* zo = (T*)Lv.z[1];  // zsa may be z[1] itself: each cell read, then written, by its thread
  const T* __restrict__ zc = (const T*)Lc.z[l + 1 == s.nlev - 1 ? 0 : 1];
  const T* zsa = addzs ? (const T*)Lv.zs : nullptr;
  const int cnx = Lc.nx, cny = Lc.ny;
  const T wa = (T)kOmegaA, wb = (T)kOmegaB;
  const long long pl = (long long)nx * ny;
  const long long row = (long long)y * nx;
  auto zw = [&](int k) { return k < 0 ? k + nz : (k >= nz ? k - nz : k); };
  // u = z[0] + P z_c at cells x0 .. x0+VX-1 of row yy, plane zz
  auto uv = [&](int yy, int zz) -> V {
    V u = ldv<T, VX>(zi, (long long)zz * pl + (long long)yy * nx + x0);
    const long long cb = ((long long)(zz >> 1) * cny + (yy >> 1)) * cnx + (x0 >> 1);
    if constexpr (VX >= 2) {
      const vec_t<T, VX / 2> c = ldv<T, VX / 2>(zc, cb);
#pragma unroll
      for (int v = 0; v < VX; ++v) u[v] += (T)kProlong * c[v / 2];
    } else {
      u[0] += (T)kProlong * zc[cb];
    }
    return u;
  };
  // sum_f k_f (w_c - w_nb) + m w_c over the x, y faces (neighbours in the LDS
  // plane sp, or from `far` for the rows outside the tile) and the z faces
  auto lap = [&](const T* sp, long long base, V w, V wzm, V wzp, V kxc, V kyc, V kzc, V kzm,
                 auto far) -> V {
    V l7 = kzc * (w - wzp) + kzm * (w - wzm);
    V wy;
    if (iy > 0) {
#pragma unroll
      for (int v = 0; v < VX; ++v) wy[v] = sp[lofs - nx + v];
    } else {
      wy = far(ym);
    }
    l7 += co.template comp<VX>(1, base + (long long)ym * nx + x0) * (w - wy);
    if (iy + 1 < ty) {
#pragma unroll
      for (int v = 0; v < VX; ++v) wy[v] = sp[lofs + nx + v];
    } else {
      wy = far(yp);
    }
    l7 += kyc * (w - wy);
    const T wxm = sp[lxm], wxp = sp[lxp];
    const T kxm = co.comp1g(0, base + row + xm);
#pragma unroll
    for (int v = 0; v < VX; ++v) {
      const T up = v + 1 < VX ? w[(v + 1) % VX] : wxp;
      const T uw = v > 0 ? w[(v + VX - 1) % VX] : wxm;
      const T kw = v > 0 ? kxc[(v + VX - 1) % VX] : kxm;
      l7[v] += kxc[v] * (w[v] - up) + kw * (w[v] - uw);
    }
    if (mm) l7 += ldv<T, VX>(mm, base + row + x0) * w;
    return l7;
  };
  // v sweep: u at planes k-1, k and kz at plane k-1; z[1] sweep: v at planes
  // k-2, k-1 (own cells), the cell's coefficients and r at plane k-1 and kz
  // at plane k-2
  V um = (T)0, uc = (T)0, kzB = (T)0;
  V vm = (T)0, vc = (T)0;
  V kx1 = (T)0, ky1 = (T)0, kz1 = (T)0, d1 = (T)0, r1 = (T)0, kzA = (T)0;
  if (active) {
    um = uv(y, zw(za - 2));
    uc = uv(y, zw(za - 1));
    kzB = co.template comp<VX>(2, (long long)zw(za - 2) * pl + row + x0);
#pragma unroll
    for (int v = 0; v < VX; ++v) su[0][lofs + v] = uc[v];
  }
  __syncthreads();
  double dot = 0.0;
  for (int k = za - 1; k <= zb; ++k) {
    const int st = k - (za - 1);
    V vn = (T)0, kxc = (T)0, kyc = (T)0, kzc = (T)0, dc = (T)0, rv = (T)0;
    // v on plane k (planes za-1 .. zb)
    if (active) {
      const int kk = zw(k);
      const V un = uv(y, zw(k + 1));
      if (k + 1 <= zb) {
#pragma unroll
        for (int v = 0; v < VX; ++v) su[(st + 1) & 1][lofs + v] = un[v];
      }
      const long long base = (long long)kk * pl;
      const long long ic = base + row + x0;
      co.template own<VX>(ic, kxc, kyc, kzc, dc);
      rv = ldv<T, VX>(r, ic);
      auto far = [&](int yy) -> V { return uv(yy, kk); };
      const V lu = lap(su[st & 1], base, uc, um, un, kxc, kyc, kzc, kzB, far);
      vn = uc + wb * dc * (rv - lu);
#pragma unroll
      for (int v = 0; v < VX; ++v) sv[st & 1][lofs + v] = vn[v];
      kzB = kzc;
      um = uc;
      uc = un;
    }
    // z[1] on plane k - 1 (output planes za .. zb-1; no wrap)
    if (k - 1 >= za && outrow) {
      const long long base = (long long)(k - 1) * pl;
      const long long ic = base + row + x0;
      auto none = [&](int) -> V { return (T)0; };  // output rows have both y neighbours
      const V lv = lap(sv[(st + 1) & 1], base, vc, vm, vn, kx1, ky1, kz1, kzA, none);
      V out = vc + wa * d1 * (r1 - lv);
      if (zsa) out += ldv<T, VX>(zsa, ic);
      stv<T, VX>(zo, ic, out);
      if constexpr (DOT) {
#pragma unroll
        for (int v = 0; v < VX; ++v) dot += (double)r1[v] * (double)out[v];
      }
    }
    vm = vc;
    vc = vn;
    kzA = kz1;
    kx1 = kxc;
    ky1 = kyc;
    kz1 = kzc;
    d1 = dc;
    r1 = rv;
    __syncthreads();
  }
  if constexpr (DOT) {
    double acc[1] = {dot};
    block_partial<1, 0u, kPostBlock>(acc, s.partials, rc.nblk, lb);
  }
}

// bfloat16 pair of a 32-bit word: element 0 in the low half
__device__ __forceinline__ float bflo(unsigned w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bfhi(unsigned w) { return __uint_as_float(w & 0xffff0000u); }
__device__ __forceinline__ vec_t<float, 4> bf4(unsigned a, unsigned b) {
  vec_t<float, 4> r;
  r[0] = bflo(a);
  r[1] = bfhi(a);
  r[2] = bflo(b);
  r[3] = bfhi(b);
  return r;
}

// k_mg_post for FP32 engines with 4 cells per thread, software-pipelined:
// the loads that go to HBM are issued one step ahead (u on plane k + 2, the
// packed coefficients, r and the halo row's u on plane k + 1), so their
// latency overlaps the barrier and the compute of a step instead of
// preceding it (the barrier waits for LDS only); the y-minus row's ky and the
// x-minus cell's kx, which neighbouring threads of the workgroup load as
// their own, come from L2 within the step.  The coefficients stay packed
// (bfloat16 pairs) in registers; the second sweep reuses the first sweep's
// ky and kx of the neighbouring cells instead of loading them again.  Same
// arithmetic as k_mg_post.
template <bool DOT, bool L0>
__global__ __launch_bounds__(kPostBlock) void k_mg_post4(const SysDev* __restrict__ sds,
                                                         const Rec* __restrict__ recs,
                                                         const int* __restrict__ b2r,
                                                         int l_arg, int wl, int so, int addzs) {
  using V = vec_t<float, 4>;
  using U4 = vec_t<unsigned, 4>;
  using U2 = vec_t<unsigned, 2>;
  using V2 = vec_t<float, 2>;
  constexpr int VX = 4;
  const int l = L0 ? 0 : l_arg;
  __shared__ float su[2][kPostBlock * VX];
  __shared__ float sv[2][kPostBlock * VX];
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const Rec rc = cload(gtab(recs + cload(gtab(b2r + b))));
  const SysDev s = cload(gtab(sds + rc.sys));
  if (s.sc->done) return;
  if (wl >= 0 && wl >= s.mg_w) return;
  if (so && !s.sc->first) return;
  const MgLevel Lv = cload(gtab(s.lev + l));
  const MgLevel Lc = cload(gtab(s.lev + l + 1));
  const int nx = Lv.nx, ny = Lv.ny, nz = Lv.nz;
  const int lb = b - rc.blk0;
  const int lx = rc.lx, ty = rc.ty;
  const int tiy = lb % rc.tiles_y, tiz = lb / rc.tiles_y;
  const int ix = threadIdx.x % lx, iy = threadIdx.x / lx;
  const bool active = iy < ty;
  const int y0 = tiy * (ty - 2);
  const int yr = y0 - 1 + iy;
  const int y = yr < 0 ? yr + ny : (yr >= ny ? yr - ny : yr);
  const bool outrow = active && iy >= 1 && iy <= ty - 2 && yr < ny;
  const int za = tiz * rc.bz, zb = min(za + rc.bz, nz);
  const int x0 = ix * VX;
  const int lofs = iy * nx + x0;
  const int lxm = iy * nx + (ix == 0 ? nx - 1 : x0 - 1);
  const int lxp = iy * nx + (ix == lx - 1 ? 0 : x0 + VX);
  const int xm = x0 == 0 ? nx - 1 : x0 - 1;
  const int ym = wrapm(y, ny), yp = wrapp(y, ny);
  const bool far = iy == 0 || iy + 1 == ty;
  const int yfar = iy == 0 ? ym : yp;
  // per-thread offsets within a plane (32-bit: a plane has < 2^31 cells);
  // plane bases are uniform over the workgroup
  const int oc = y * nx + x0;                     // the thread's cells
  const int oym = ym * nx + x0;                   // the y-minus row
  const int oxm = y * nx + xm;                    // the x-minus cell
  const int ofar = yfar * nx + x0;                // the halo row (edge thread rows)
  const int cnx = Lc.nx, cny = Lc.ny;
  const int occ = (y >> 1) * cnx + (x0 >> 1);     // coarse cells under oc
  const int ocf = (yfar >> 1) * cnx + (x0 >> 1);  // coarse cells under ofar
  const long long pl = (long long)nx * ny;
  const long long cpl = (long long)cnx * cny;
  const gcp<unsigned short> kc = (gcp<unsigned short>)Lv.kc;
  const gcp<float> mm = (gcp<float>)Lv.m;
  const gcp<float> r = (gcp<float>)Lv.r;
  const gcp<float> zi = (gcp<float>)Lv.z[0];
  const gmp<float> zo = (gmp<float>)Lv.z[1];
  const gcp<float> zc = (gcp<float>)Lc.z[l + 1 == s.nlev - 1 ? 0 : 1];
  const gcp<float> zsa = addzs ? (gcp<float>)Lv.zs : nullptr;
  const float wa = (float)kOmegaA, wb = (float)kOmegaB;
  auto zw = [&](int k) { return k < 0 ? k + nz : (k >= nz ? k - nz : k); };
  // u = z[0] + P z_c at the 4 cells at offset o (coarse offset co) of plane zz
  auto uv = [&](int zz, int o, int co) -> V {
    V u = *(gcp<V>)(zi + (long long)zz * pl + o);
    const V2 c = *(gcp<V2>)(zc + (long long)(zz >> 1) * cpl + co);
#pragma unroll
    for (int v = 0; v < VX; ++v) u[v] += (float)kProlong * c[v / 2];
    return u;
  };
  // the packed coefficients of plane kk: group of the cells at offset o
  // (o % 4 == 0) at kcp + 4 o
  auto kplane = [&](int kk) { return kc + 4 * (long long)kk * pl; };
  // the laplacian of k_mg_post with ky of the y-minus row and kx of the
  // x-minus cell given
  auto lap = [&](const float* sp, long long base, V w, V wzm, V wzp, V kxc, V kyc, V kzc,
                 V kzm, V kym, float kxm, V wfar) -> V {
    V l7 = kzc * (w - wzp) + kzm * (w - wzm);
    V wy;
    if (iy > 0) {
#pragma unroll
      for (int v = 0; v < VX; ++v) wy[v] = sp[lofs - nx + v];
    } else {
      wy = wfar;
    }
    l7 += kym * (w - wy);
    if (iy + 1 < ty) {
#pragma unroll
      for (int v = 0; v < VX; ++v) wy[v] = sp[lofs + nx + v];
    } else {
      wy = wfar;
    }
    l7 += kyc * (w - wy);
    const float wxm = sp[lxm], wxp = sp[lxp];
#pragma unroll
    for (int v = 0; v < VX; ++v) {
      const float up = v + 1 < VX ? w[(v + 1) % VX] : wxp;
      const float uw = v > 0 ? w[(v + VX - 1) % VX] : wxm;
      const float kw = v > 0 ? kxc[(v + VX - 1) % VX] : kxm;
      l7[v] += kxc[v] * (w[v] - up) + kw * (w[v] - uw);
    }
    if (mm) l7 += *(gcp<V>)(mm + base + oc) * w;
    return l7;
  };
  // prefetched operands of the first sweep on plane k: u on plane k + 1,
  // the packed coefficients, r, ky of row ym, kx of cell xm and the halo
  // row's u (first and last thread rows) on plane k
  V pu = 0.0f, pr = 0.0f, pfar = 0.0f;
  U4 pk0 = 0u, pk1 = 0u;
  auto prefetch = [&](int k) {
    const int kk = zw(k);
    const gcp<unsigned short> kp = kplane(kk);
    pu = uv(zw(k + 1), oc, occ);
    const gcp<U4> g = (gcp<U4>)(kp + 4 * oc);
    pk0 = g[0];
    pk1 = g[1];
    pr = *(gcp<V>)(r + (long long)kk * pl + oc);
    if (far) pfar = uv(kk, ofar, ocf);
  };
  // state: u on planes k-1 (um), k (uc); v on planes k-2 (vm), k-1 (vc);
  // packed kz of planes k-1 (zB) and k-2 (zA); plane k-1's packed
  // coefficients, r, ky of row ym and kx of cell xm for the second sweep
  V um = 0.0f, uc = 0.0f, vm = 0.0f, vc = 0.0f, r1 = 0.0f;
  U2 zA = 0u, zB = 0u, ky1 = 0u;
  U4 q0 = 0u, q1 = 0u;
  unsigned short kx1 = 0;
  if (active) {
    um = uv(zw(za - 2), oc, occ);
    uc = uv(zw(za - 1), oc, occ);
    zB = *(gcp<U2>)(kplane(zw(za - 2)) + 4 * oc + 8);  // kz[4]
#pragma unroll
    for (int v = 0; v < VX; ++v) su[0][lofs + v] = uc[v];
    prefetch(za - 1);
  }
  __syncthreads();
  double dot = 0.0;
  for (int k = za - 1; k <= zb; ++k) {
    const int st = k - (za - 1);
    // this step's operands (loaded during the previous step)
    const V un = pu, rv = pr, wfar = pfar;
    const U4 c0 = pk0, c1 = pk1;
    U2 cy = 0u;
    unsigned short cx = 0;
    if (active) {
      const gcp<unsigned short> kp = kplane(zw(k));
      cy = *(gcp<U2>)(kp + 4 * oym + 4);
      cx = kp[((oxm >> 2) << 4) + (oxm & 3)];
    }
    if (active && k + 1 <= zb) prefetch(k + 1);
    V vn = 0.0f;
    // v on plane k (planes za-1 .. zb)
    if (active) {
      const int kk = zw(k);
      if (k + 1 <= zb) {
#pragma unroll
        for (int v = 0; v < VX; ++v) su[(st + 1) & 1][lofs + v] = un[v];
      }
      const V lu = lap(su[st & 1], (long long)kk * pl, uc, um, un, bf4(c0.x, c0.y),
                       bf4(c0.z, c0.w), bf4(c1.x, c1.y), bf4(zB.x, zB.y), bf4(cy.x, cy.y),
                       bflo(cx), wfar);
      vn = uc + wb * bf4(c1.z, c1.w) * (rv - lu);
#pragma unroll
      for (int v = 0; v < VX; ++v) sv[st & 1][lofs + v] = vn[v];
      um = uc;
      uc = un;
    }
    // z[1] on plane k - 1 (output planes za .. zb-1; no wrap)
    if (k - 1 >= za && outrow) {
      const long long base = (long long)(k - 1) * pl;
      const V lv = lap(sv[(st + 1) & 1], base, vc, vm, vn, bf4(q0.x, q0.y), bf4(q0.z, q0.w),
                       bf4(q1.x, q1.y), bf4(zA.x, zA.y), bf4(ky1.x, ky1.y), bflo(kx1),
                       V(0.0f));
      V out = vc + wa * bf4(q1.z, q1.w) * (r1 - lv);
      if (zsa) out += *(gcp<V>)(zsa + base + oc);
      *(gmp<V>)(zo + base + oc) = out;
      if constexpr (DOT) {
#pragma unroll
        for (int v = 0; v < VX; ++v) dot += (double)r1[v] * (double)out[v];
      }
    }
    vm = vc;
    vc = vn;
    zA = zB;
    zB = U2{c1.x, c1.y};
    q0 = c0;
    q1 = c1;
    r1 = rv;
    ky1 = cy;
    kx1 = cx;
    __syncthreads();
  }
  if constexpr (DOT) {
    double acc[1] = {dot};
    block_partial<1, 0u, kPostBlock>(acc, s.partials, rc.nblk, lb);
  }
}

// Pre-smoothing and restriction of a level in one z march (the levels with
// post = 1, on the tiles of k_mg_post: an even number of output rows from an
// even row, an even number of planes from an even plane):
//   u = wA D^-1 r;  z = u + wB D^-1 (r - L u) -> z[0] (output rows and planes);
//   r_{l+1}[C] = sum over the 2x2x2 children c of (r - L z)_c.
// Step k forms z on plane k (halo rows and the planes za-1, zb included) and
// then the residual on plane k-1 (plane neighbours of z in LDS since the
// previous step; the cell's own z on planes k-2 .. k and its coefficients and
// r on plane k-1 in registers).  x pairs are summed in the thread, plane
// pairs across two steps, row pairs through LDS after the next barrier (odd
// rows hand their sums to the even row above).  One barrier per step; u and z
// planes double-buffered in LDS.
template <typename T, int VX, bool L0>
__global__ __launch_bounds__(kPostBlock) void k_mg_pre(const SysDev* __restrict__ sds,
                                                       const Rec* __restrict__ recs,
                                                       const int* __restrict__ b2r, int l_arg,
                                                       int wl, int so) {
  static_assert(VX >= 2, "x pairs of the restriction are summed in the thread");
  constexpr int NP = VX / 2;
  const int l = L0 ? 0 : l_arg;
  using V = vec_t<T, VX>;
  __shared__ T su[2][kPostBlock * VX];
  __shared__ T sz[2][kPostBlock * VX];
  __shared__ T sacc[kPostBlock * NP];
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  // launch tables (recs, b2r, sds) and level tables (s.lev) are device
  // arrays written by the host between launches only
  const Rec rc = cload(gtab(recs + cload(gtab(b2r + b))));
  const SysDev s = cload(gtab(sds + rc.sys));
  if (s.sc->done) return;
  if (wl >= 0 && wl >= s.mg_w) return;
  if (so && !s.sc->first) return;
  const MgLevel Lv = cload(gtab(s.lev + l));
  const MgLevel Lc = cload(gtab(s.lev + l + 1));
  const int nx = Lv.nx, ny = Lv.ny, nz = Lv.nz;
  const int lb = b - rc.blk0;
  const int lx = rc.lx, ty = rc.ty;  // lx * VX == nx, ty even, <= kPostBlock / lx
  const int tiy = lb % rc.tiles_y, tiz = lb / rc.tiles_y;
  const int ix = threadIdx.x % lx, iy = threadIdx.x / lx;
  const bool active = iy < ty;
  const int y0 = tiy * (ty - 2);  // even
  const int yr = y0 - 1 + iy;
  const int y = yr < 0 ? yr + ny : (yr >= ny ? yr - ny : yr);
  const bool outrow = active && iy >= 1 && iy <= ty - 2 && yr < ny;
  const int za = tiz * rc.bz, zb = min(za + rc.bz, nz);  // za, zb even
  const int x0 = ix * VX;
  const int lofs = iy * nx + x0;
  const int lxm = iy * nx + (ix == 0 ? nx - 1 : x0 - 1);
  const int lxp = iy * nx + (ix == lx - 1 ? 0 : x0 + VX);
  const int xm = x0 == 0 ? nx - 1 : x0 - 1;
  const int ym = wrapm(y, ny), yp = wrapp(y, ny);
  const MgCoefs<T> co(Lv);
  const T* __restrict__ mm = (const T*)Lv.m;
  const T* __restrict__ r = (const T*)Lv.r;
  T* __restrict__ zo = (T*)Lv.z[0];
  T* __restrict__ rco = (T*)Lc.r;
  const int cnx = Lc.nx, cny = Lc.ny;
  const T wa = (T)kOmegaA, wb = (T)kOmegaB;
  const long long pl = (long long)nx * ny;
  const long long row = (long long)y * nx;
  auto zw = [&](int k) { return k < 0 ? k + nz : (k >= nz ? k - nz : k); };
  // u = wA D^-1 r (the zero-guess sweep) at row yy, plane zz
  auto uv = [&](int yy, int zz) -> V {
    const long long i = (long long)zz * pl + (long long)yy * nx + x0;
    return wa * co.template comp<VX>(3, i) * ldv<T, VX>(r, i);
  };
  auto lap = [&](const T* sp, long long base, V w, V wzm, V wzp, V kxc, V kyc, V kzc, V kzm,
                 auto far) -> V {
    V l7 = kzc * (w - wzp) + kzm * (w - wzm);
    V wy;
    if (iy > 0) {
#pragma unroll
      for (int v = 0; v < VX; ++v) wy[v] = sp[lofs - nx + v];
    } else {
      wy = far(ym);
    }
    l7 += co.template comp<VX>(1, base + (long long)ym * nx + x0) * (w - wy);
    if (iy + 1 < ty) {
#pragma unroll
      for (int v = 0; v < VX; ++v) wy[v] = sp[lofs + nx + v];
    } else {
      wy = far(yp);
    }
    l7 += kyc * (w - wy);
    const T wxm = sp[lxm], wxp = sp[lxp];
    const T kxm = co.comp1g(0, base + row + xm);
#pragma unroll
    for (int v = 0; v < VX; ++v) {
      const T up = v + 1 < VX ? w[(v + 1) % VX] : wxp;
      const T uw = v > 0 ? w[(v + VX - 1) % VX] : wxm;
      const T kw = v > 0 ? kxc[(v + VX - 1) % VX] : kxm;
      l7[v] += kxc[v] * (w[v] - up) + kw * (w[v] - uw);
    }
    if (mm) l7 += ldv<T, VX>(mm, base + row + x0) * w;
    return l7;
  };
  // z sweep: u at planes k-1, k, r at plane k, kz at plane k-1; residual: z at
  // planes k-2, k-1 (own cells), the cell's coefficients and r at plane k-1,
  // kz at plane k-2; the x-pair sums of the current plane pair; even rows
  // keep their sums of the last pair for the row pair
  V um = (T)0, uc = (T)0, rcur = (T)0, kzB = (T)0;
  V zm = (T)0, zc0 = (T)0;
  V kx1 = (T)0, ky1 = (T)0, kz1 = (T)0, r1 = (T)0, kzA = (T)0;
  T racc[NP], rpend[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) racc[j] = rpend[j] = (T)0;
  int kpend = -1;  // coarse plane of rpend (even rows), -1: none
  if (active) {
    um = uv(y, zw(za - 2));
    const long long i0 = (long long)zw(za - 1) * pl + row + x0;
    rcur = ldv<T, VX>(r, i0);
    uc = wa * co.template comp<VX>(3, i0) * rcur;
    kzB = co.template comp<VX>(2, (long long)zw(za - 2) * pl + row + x0);
#pragma unroll
    for (int v = 0; v < VX; ++v) su[0][lofs + v] = uc[v];
  }
  __syncthreads();
  for (int k = za - 1; k <= zb; ++k) {
    const int st = k - (za - 1);
    // row pairs of the plane pair completed in the previous step
    if (kpend >= 0) {
#pragma unroll
      for (int j = 0; j < NP; ++j)
        st1(rco, ((long long)kpend * cny + (y >> 1)) * cnx + (x0 >> 1) + j,
            rpend[j] + sacc[(threadIdx.x + lx) * NP + j]);
      kpend = -1;
    }
    V zn = (T)0, kxc = (T)0, kyc = (T)0, kzc = (T)0, rk = (T)0;
    // z on plane k (planes za-1 .. zb)
    if (active) {
      const int kk = zw(k);
      const long long in = (long long)zw(k + 1) * pl + row + x0;
      const V rn = ldv<T, VX>(r, in);
      const V un = wa * co.template comp<VX>(3, in) * rn;
      if (k + 1 <= zb) {
#pragma unroll
        for (int v = 0; v < VX; ++v) su[(st + 1) & 1][lofs + v] = un[v];
      }
      const long long base = (long long)kk * pl;
      const long long ic = base + row + x0;
      V dc;
      co.template own<VX>(ic, kxc, kyc, kzc, dc);
      auto far = [&](int yy) -> V { return uv(yy, kk); };
      const V lu = lap(su[st & 1], base, uc, um, un, kxc, kyc, kzc, kzB, far);
      zn = uc + wb * dc * (rcur - lu);
#pragma unroll
      for (int v = 0; v < VX; ++v) sz[st & 1][lofs + v] = zn[v];
      if (outrow && k >= za && k < zb) stv<T, VX>(zo, ic, zn);
      rk = rcur;
      kzB = kzc;
      um = uc;
      uc = un;
      rcur = rn;
    }
    // residual on plane k - 1 (planes za .. zb-1, no wrap)
    if (k - 1 >= za && outrow) {
      const int kp = k - 1;
      const long long base = (long long)kp * pl;
      auto none = [&](int) -> V { return (T)0; };
      const V lz = lap(sz[(st + 1) & 1], base, zc0, zm, zn, kx1, ky1, kz1, kzA, none);
      const V res = r1 - lz;
      if ((kp & 1) == 0) {
#pragma unroll
        for (int j = 0; j < NP; ++j) racc[j] = res[2 * j] + res[2 * j + 1];
      } else {
#pragma unroll
        for (int j = 0; j < NP; ++j) racc[j] += res[2 * j] + res[2 * j + 1];
        if (y & 1) {
#pragma unroll
          for (int j = 0; j < NP; ++j) sacc[threadIdx.x * NP + j] = racc[j];
        } else {
#pragma unroll
          for (int j = 0; j < NP; ++j) rpend[j] = racc[j];
          kpend = kp >> 1;
        }
      }
    }
    zm = zc0;
    zc0 = zn;
    kzA = kz1;
    kx1 = kxc;
    ky1 = kyc;
    kz1 = kzc;
    r1 = rk;
    __syncthreads();
  }
  // the last plane pair (planes zb-2, zb-1)
  if (kpend >= 0) {
#pragma unroll
    for (int j = 0; j < NP; ++j)
      st1(rco, ((long long)kpend * cny + (y >> 1)) * cnx + (x0 >> 1) + j,
          rpend[j] + sacc[(threadIdx.x + lx) * NP + j]);
  }
}

// k_mg_pre for FP32 engines with 4 cells per thread, software-pipelined like
// k_mg_post4: the HBM loads of step k (r and D^-1 on plane k + 2, the packed
// coefficients and the halo row's r and D^-1 on plane k + 1) are for step
// k + 1; plane k-1's packed coefficients, r, ky and kx wait in LDS slots for
// the residual.  Same arithmetic as k_mg_pre.
template <bool L0>
__global__ __launch_bounds__(kPostBlock) void k_mg_pre4(const SysDev* __restrict__ sds,
                                                        const Rec* __restrict__ recs,
                                                        const int* __restrict__ b2r, int l_arg,
                                                        int wl, int so) {
  using V = vec_t<float, 4>;
  using U4 = vec_t<unsigned, 4>;
  using U2 = vec_t<unsigned, 2>;
  constexpr int VX = 4;
  constexpr int NP = VX / 2;
  const int l = L0 ? 0 : l_arg;
  __shared__ float su[2][kPostBlock * VX];
  __shared__ float sz[2][kPostBlock * VX];
  __shared__ float sacc[kPostBlock * NP];
  // each thread's plane k-1 data for the residual (own slots only): packed
  // coefficients, r, and {ky of row ym, kz of plane k-2}
  __shared__ U4 sq[2][kPostBlock];
  __shared__ V sr[kPostBlock];
  __shared__ U4 sm[kPostBlock];
  __shared__ unsigned skx[kPostBlock];               // kx of cell xm, plane k-1
  __shared__ vec_t<float, NP> spend[kPostBlock];     // even rows' pending sums
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const Rec rc = cload(gtab(recs + cload(gtab(b2r + b))));
  const SysDev s = cload(gtab(sds + rc.sys));
  if (s.sc->done) return;
  if (wl >= 0 && wl >= s.mg_w) return;
  if (so && !s.sc->first) return;
  const MgLevel Lv = cload(gtab(s.lev + l));
  const MgLevel Lc = cload(gtab(s.lev + l + 1));
  const int nx = Lv.nx, ny = Lv.ny, nz = Lv.nz;
  const int lb = b - rc.blk0;
  const int lx = rc.lx, ty = rc.ty;
  const int tiy = lb % rc.tiles_y, tiz = lb / rc.tiles_y;
  const int ix = threadIdx.x % lx, iy = threadIdx.x / lx;
  const bool active = iy < ty;
  const int y0 = tiy * (ty - 2);
  const int yr = y0 - 1 + iy;
  const int y = yr < 0 ? yr + ny : (yr >= ny ? yr - ny : yr);
  const bool outrow = active && iy >= 1 && iy <= ty - 2 && yr < ny;
  const int za = tiz * rc.bz, zb = min(za + rc.bz, nz);
  const int x0 = ix * VX;
  const int lofs = iy * nx + x0;
  const int lxm = iy * nx + (ix == 0 ? nx - 1 : x0 - 1);
  const int lxp = iy * nx + (ix == lx - 1 ? 0 : x0 + VX);
  const int xm = x0 == 0 ? nx - 1 : x0 - 1;
  const int ym = wrapm(y, ny), yp = wrapp(y, ny);
  const bool far = iy == 0 || iy + 1 == ty;
  const int yfar = iy == 0 ? ym : yp;
  const int oc = y * nx + x0;
  const int oym = ym * nx + x0;
  const int oxm = y * nx + xm;
  const int ofar = yfar * nx + x0;
  const long long pl = (long long)nx * ny;
  const gcp<unsigned short> kc = (gcp<unsigned short>)Lv.kc;
  const gcp<float> mm = (gcp<float>)Lv.m;
  const gcp<float> r = (gcp<float>)Lv.r;
  const gmp<float> zo = (gmp<float>)Lv.z[0];
  const gmp<float> rco = (gmp<float>)Lc.r;
  const int cnx = Lc.nx, cny = Lc.ny;
  const float wa = (float)kOmegaA, wb = (float)kOmegaB;
  auto zw = [&](int k) { return k < 0 ? k + nz : (k >= nz ? k - nz : k); };
  auto kplane = [&](int kk) { return kc + 4 * (long long)kk * pl; };
  // u = wA D^-1 r from r and the packed D^-1 pair of 4 cells
  auto u_of = [&](V rr, U2 d) -> V { return wa * bf4(d.x, d.y) * rr; };
  auto lap = [&](const float* sp, long long base, V w, V wzm, V wzp, V kxc, V kyc, V kzc,
                 V kzm, V kym, float kxm, V wfar) -> V {
    V l7 = kzc * (w - wzp) + kzm * (w - wzm);
    V wy;
    if (iy > 0) {
#pragma unroll
      for (int v = 0; v < VX; ++v) wy[v] = sp[lofs - nx + v];
    } else {
      wy = wfar;
    }
    l7 += kym * (w - wy);
    if (iy + 1 < ty) {
#pragma unroll
      for (int v = 0; v < VX; ++v) wy[v] = sp[lofs + nx + v];
    } else {
      wy = wfar;
    }
    l7 += kyc * (w - wy);
    const float wxm = sp[lxm], wxp = sp[lxp];
#pragma unroll
    for (int v = 0; v < VX; ++v) {
      const float up = v + 1 < VX ? w[(v + 1) % VX] : wxp;
      const float uw = v > 0 ? w[(v + VX - 1) % VX] : wxm;
      const float kw = v > 0 ? kxc[(v + VX - 1) % VX] : kxm;
      l7[v] += kxc[v] * (w[v] - up) + kw * (w[v] - uw);
    }
    if (mm) l7 += *(gcp<V>)(mm + base + oc) * w;
    return l7;
  };
  // prefetched operands of the z sweep on plane k: r and D^-1 on plane k + 1
  // (u there, and r of the next step), the packed coefficients, ky of row
  // ym, kx of cell xm and the halo row's r and D^-1 on plane k
  V pr = 0.0f, pfr = 0.0f;
  U2 pd = 0u, pfd = 0u;
  U4 pk0 = 0u, pk1 = 0u;
  auto prefetch = [&](int k) {
    const int kk = zw(k), kn = zw(k + 1);
    const gcp<unsigned short> kp = kplane(kk);
    pr = *(gcp<V>)(r + (long long)kn * pl + oc);
    pd = *(gcp<U2>)(kplane(kn) + 4 * oc + 12);  // dinv[4]: the last quarter
    const gcp<U4> g = (gcp<U4>)(kp + 4 * oc);
    pk0 = g[0];
    pk1 = g[1];
    if (far) {
      pfr = *(gcp<V>)(r + (long long)kk * pl + ofar);
      pfd = *(gcp<U2>)(kp + 4 * ofar + 12);
    }
  };
  // state: u on planes k-1 (um), k (uc), r on plane k (rcur); z on planes
  // k-2 (zm), k-1 (zc0); packed kz of planes k-1 (zB), k-2 (zA); r, ky of
  // row ym and kx of cell xm on plane k-1 (r1, ky1, kx1; its packed
  // coefficients in sq); the restriction sums
  V um = 0.0f, uc = 0.0f, rcur = 0.0f, zm = 0.0f, zc0 = 0.0f;
  U2 zB = 0u;
  float racc[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) racc[j] = 0.0f;
  int kpend = -1;  // coarse plane of rpend (even rows), -1: none
  if (active) {
    const int k2 = zw(za - 2), k1 = zw(za - 1);
    um = u_of(*(gcp<V>)(r + (long long)k2 * pl + oc), *(gcp<U2>)(kplane(k2) + 4 * oc + 12));
    rcur = *(gcp<V>)(r + (long long)k1 * pl + oc);
    uc = u_of(rcur, *(gcp<U2>)(kplane(k1) + 4 * oc + 12));
    zB = *(gcp<U2>)(kplane(k2) + 4 * oc + 8);  // kz[4]
#pragma unroll
    for (int v = 0; v < VX; ++v) su[0][lofs + v] = uc[v];
    prefetch(za - 1);
  }
  __syncthreads();
  for (int k = za - 1; k <= zb; ++k) {
    const int st = k - (za - 1);
    // row pairs of the plane pair completed in the previous step
    if (kpend >= 0) {
      const vec_t<float, NP> rp = spend[threadIdx.x];
#pragma unroll
      for (int j = 0; j < NP; ++j)
        rco[((long long)kpend * cny + (y >> 1)) * cnx + (x0 >> 1) + j] =
            rp[j] + sacc[(threadIdx.x + lx) * NP + j];
      kpend = -1;
    }
    const V rn = pr, fr = pfr;
    const U2 dn = pd, fd = pfd;
    const U4 c0 = pk0, c1 = pk1;
    U2 cy = 0u;
    unsigned short cx = 0;
    if (active) {
      const gcp<unsigned short> kp = kplane(zw(k));
      cy = *(gcp<U2>)(kp + 4 * oym + 4);  // ky[4] of the y-minus row
      cx = kp[((oxm >> 2) << 4) + (oxm & 3)];
    }
    if (active && k + 1 <= zb) prefetch(k + 1);
    V zn = 0.0f, rk = 0.0f;
    // z on plane k (planes za-1 .. zb)
    if (active) {
      const int kk = zw(k);
      const V un = u_of(rn, dn);
      if (k + 1 <= zb) {
#pragma unroll
        for (int v = 0; v < VX; ++v) su[(st + 1) & 1][lofs + v] = un[v];
      }
      const long long base = (long long)kk * pl;
      const V lu = lap(su[st & 1], base, uc, um, un, bf4(c0.x, c0.y), bf4(c0.z, c0.w),
                       bf4(c1.x, c1.y), bf4(zB.x, zB.y), bf4(cy.x, cy.y), bflo(cx),
                       u_of(fr, fd));
      zn = uc + wb * bf4(c1.z, c1.w) * (rcur - lu);
#pragma unroll
      for (int v = 0; v < VX; ++v) sz[st & 1][lofs + v] = zn[v];
      if (outrow && k >= za && k < zb) *(gmp<V>)(zo + base + oc) = zn;
      rk = rcur;
      um = uc;
      uc = un;
      rcur = rn;
    }
    // residual on plane k - 1 (planes za .. zb-1, no wrap)
    if (k - 1 >= za && outrow) {
      const int kp = k - 1;
      const long long base = (long long)kp * pl;
      const U4 q0 = sq[0][threadIdx.x], q1 = sq[1][threadIdx.x], qm = sm[threadIdx.x];
      const V lz = lap(sz[(st + 1) & 1], base, zc0, zm, zn, bf4(q0.x, q0.y), bf4(q0.z, q0.w),
                       bf4(q1.x, q1.y), bf4(qm.z, qm.w), bf4(qm.x, qm.y), bflo(skx[threadIdx.x]),
                       V(0.0f));
      const V res = sr[threadIdx.x] - lz;
      if ((kp & 1) == 0) {
#pragma unroll
        for (int j = 0; j < NP; ++j) racc[j] = res[2 * j] + res[2 * j + 1];
      } else {
#pragma unroll
        for (int j = 0; j < NP; ++j) racc[j] += res[2 * j] + res[2 * j + 1];
        if (y & 1) {
#pragma unroll
          for (int j = 0; j < NP; ++j) sacc[threadIdx.x * NP + j] = racc[j];
        } else {
          vec_t<float, NP> rp;
#pragma unroll
          for (int j = 0; j < NP; ++j) rp[j] = racc[j];
          spend[threadIdx.x] = rp;
          kpend = kp >> 1;
        }
      }
    }
    zm = zc0;
    zc0 = zn;
    sm[threadIdx.x] = U4{cy.x, cy.y, zB.x, zB.y};
    zB = U2{c1.x, c1.y};
    sq[0][threadIdx.x] = c0;
    sq[1][threadIdx.x] = c1;
    sr[threadIdx.x] = rk;
    skx[threadIdx.x] = cx;
    __syncthreads();
  }
  // the last plane pair (planes zb-2, zb-1)
  if (kpend >= 0) {
    const vec_t<float, NP> rp = spend[threadIdx.x];
#pragma unroll
    for (int j = 0; j < NP; ++j)
      rco[((long long)kpend * cny + (y >> 1)) * cnx + (x0 >> 1) + j] =
          rp[j] + sacc[(threadIdx.x + lx) * NP + j];
  }
}

// A level of at most 64 cells held by the lanes of one wavefront (cell i on
// lane i): neighbour values by lane shuffles, no LDS and no barriers.  Lanes
// past the level's cells carry zero coefficients and read themselves.
template <typename T>
struct WaveLevel {
  int n;
  int nb[6];  // lanes of the xm, xp, ym, yp, zm, zp neighbours
  T kx, kxm, ky, kym, kz, kzm, d, m;
  __device__ void load(const MgLevel& L, int lane) {
    n = (int)L.n;
    const bool on = lane < n;
    const int c = on ? lane : 0;
    const int x = c % L.nx, y = (c / L.nx) % L.ny, z = c / (L.nx * L.ny);
    const int pl = L.nx * L.ny;
    nb[0] = on ? c - x + wrapm(x, L.nx) : lane;
    nb[1] = on ? c - x + wrapp(x, L.nx) : lane;
    nb[2] = on ? c + (wrapm(y, L.ny) - y) * L.nx : lane;
    nb[3] = on ? c + (wrapp(y, L.ny) - y) * L.nx : lane;
    nb[4] = on ? c + (wrapm(z, L.nz) - z) * pl : lane;
    nb[5] = on ? c + (wrapp(z, L.nz) - z) * pl : lane;
    const MgCoefs<T> co(L);
    kx = on ? co.comp1(0, c) : (T)0;
    kxm = on ? co.comp1(0, nb[0]) : (T)0;
    ky = on ? co.comp1(1, c) : (T)0;
    kym = on ? co.comp1(1, nb[2]) : (T)0;
    kz = on ? co.comp1(2, c) : (T)0;
    kzm = on ? co.comp1(2, nb[4]) : (T)0;
    d = on ? co.comp1(3, c) : (T)0;
    m = (on && L.m) ? ((const T*)L.m)[c] : (T)0;
  }
  // (L u) of this lane's cell; every lane of the wavefront must call it
  __device__ T lap(T u) const {
    const T uxm = __shfl(u, nb[0], 64), uxp = __shfl(u, nb[1], 64);
    const T uym = __shfl(u, nb[2], 64), uyp = __shfl(u, nb[3], 64);
    const T uzm = __shfl(u, nb[4], 64), uzp = __shfl(u, nb[5], 64);
    T s = kx * (u - uxp) + kxm * (u - uxm) + ky * (u - uyp) + kym * (u - uym) + kz * (u - uzp) +
          kzm * (u - uzm);
    return s + m * u;
  }
};

// The bottom of the cycle: levels lb..top of a system (its cycle from level
// lb down; lb > every W level) in one workgroup.  Levels of more than 64
// cells: z and a second buffer of every level in LDS and, when they fit
// (bottom_full_bytes), the levels' coefficients, mass and right-hand sides
// too (copied in once; the restriction then writes the coarse right-hand
// sides to LDS), barriers between the steps; the last one or two levels of at
// most 64 cells: one wavefront, a cell per lane.  The same operations as the
// per-level launches: zero-guess sweep + sweep, restriction, `sweeps` damped
// Jacobi sweeps from zero on the coarsest level, prolongation + sweep + sweep;
// the correction of level lb goes to its z[1] (z[0] when lb is the coarsest).
// Each system takes part only in the launch for its own lb (s.mg_lb == l0).
// The engine chooses lb so that the levels fit in LDS.
constexpr int kBottomThreads = 1024;

template <typename T>
__global__ __launch_bounds__(kBottomThreads) void k_mg_bottom(const SysDev* __restrict__ sds,
                                                              int sweeps, int l0, int wl,
                                                              int so) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ MgLevel lv[kMaxLevels];  // the levels, with LDS pointers if resident
  const SysDev& s = sds[blockIdx.x];
  if (s.nlev <= 0 || s.sc->done || s.mg_lb != l0) return;
  if (wl >= 0 && wl >= s.mg_w) return;
  if (so && !s.sc->first) return;
  const int top = s.nlev - 1;
  const int tid = threadIdx.x, nt = blockDim.x;
  const T w = (T)kOmega, wa = (T)kOmegaA, wb = (T)kOmegaB;
  constexpr int es = sizeof(T);
  long long full_bytes = 0;
  for (int l = l0; l <= top; ++l) full_bytes += bottom_full_bytes(s.lev[l].n, es, s.has_m);
  const bool full = full_bytes <= kBottomLdsMax;
  // z and the second buffer of each level (pointer tables in LDS: indexed
  // by level at run time, they would otherwise live in scratch)
  __shared__ T* za[kMaxLevels];
  __shared__ T* zb[kMaxLevels];
  unsigned char* p = smem;
  for (int l = l0; l <= top; ++l) {
    if (tid == 0) za[l] = (T*)p;
    p += lds_al16(s.lev[l].n * es);
    if (tid == 0) zb[l] = (T*)p;
    p += lds_al16(s.lev[l].n * es);
  }
  if (tid == 0) {
    for (int l = l0; l <= top; ++l) lv[l] = s.lev[l];
  }
  // the copies into LDS as one table of segments (16-byte units; a tail of
  // 4-byte words per segment), so that every thread issues all of its loads
  // before it waits for any
  constexpr int kSeg = 6 * kMaxLevels;
  __shared__ const uint4* cs_src[kSeg];
  __shared__ uint4* cs_dst[kSeg];
  __shared__ int cs_beg[kSeg + 1];
  __shared__ int cs_tail[kSeg];
  int nseg = 0;
  if (full) {
    // record a copy of `bytes` (a multiple of 4) from global memory to LDS
    auto copy = [&](void* dst, const void* src, long long bytes) {
      if (tid == 0) {
        cs_src[nseg] = (const uint4*)src;
        cs_dst[nseg] = (uint4*)dst;
        cs_beg[nseg + 1] = cs_beg[nseg] + (int)(bytes >> 4);
        cs_tail[nseg] = (int)((bytes & 15) >> 2);
      }
      ++nseg;
    };
    if (tid == 0) cs_beg[0] = 0;
    for (int l = l0; l <= top; ++l) {
      const MgLevel& G = s.lev[l];
      const long long n = G.n;
      if constexpr (es == 4) {
        copy(p, G.kc, bottom_coef_bytes(n, es));
        if (tid == 0) lv[l].kc = p;
        p += lds_al16(bottom_coef_bytes(n, es));
      } else {
        const void* src[4] = {G.kx, G.ky, G.kz, G.dinv};
        for (int c = 0; c < 4; ++c) {
          copy(p, src[c], 8 * n);
          if (tid == 0) {
            if (c == 0) lv[l].kx = p;
            if (c == 1) lv[l].ky = p;
            if (c == 2) lv[l].kz = p;
            if (c == 3) lv[l].dinv = p;
          }
          p += lds_al16(8 * n);
        }
      }
      if (s.has_m) {
        copy(p, G.m, n * es);
        if (tid == 0) lv[l].m = p;
        p += lds_al16(n * es);
      }
      // right-hand side: level l0's from memory, the coarser ones restricted here
      if (l == l0) copy(p, G.r, n * es);
      if (tid == 0) lv[l].r = p;
      p += lds_al16(n * es);
    }
    __syncthreads();
    const int total = cs_beg[nseg];
    using W4 = vec_t<unsigned, 4>;
    for (int b0 = tid; b0 < total; b0 += 4 * nt) {
      W4 v0 = 0u, v1 = 0u, v2 = 0u, v3 = 0u;
      int s0 = 0, s1 = 0, s2 = 0, s3 = 0;
      auto load = [&](int u, int& sg, W4& v) {
        if (u < total) {
          while (cs_beg[sg + 1] <= u) ++sg;
          v = *(glb_cp<W4>)((const W4*)cs_src[sg] + (u - cs_beg[sg]));
        }
      };
      auto store = [&](int u, int sg, W4 v) {
        if (u < total) *(lds_p<W4>)((W4*)cs_dst[sg] + (u - cs_beg[sg])) = v;
      };
      load(b0, s0, v0);
      load(b0 + nt, s1, v1);
      load(b0 + 2 * nt, s2, v2);
      load(b0 + 3 * nt, s3, v3);
      store(b0, s0, v0);
      store(b0 + nt, s1, v1);
      store(b0 + 2 * nt, s2, v2);
      store(b0 + 3 * nt, s3, v3);
    }
    for (int k = 0; k < nseg; ++k) {
      const int n16 = cs_beg[k + 1] - cs_beg[k];
      if (tid < cs_tail[k])
        ((unsigned*)(cs_dst[k] + n16))[tid] = ((const unsigned*)(cs_src[k] + n16))[tid];
    }
  }
  __syncthreads();
  // levels lw..top go to one wavefront (a level below one of <= 64 cells has
  // <= 8, and a coarser one would have < 8: at most two such levels)
  int lw = l0;
  while (lw <= top && lv[lw].n > 64) ++lw;
  // out = u + om D^-1 (r - L u) on level l, u and out in LDS
  // f(view) with the level's data in LDS or in global memory
  auto with_view = [&](int l, auto f) {
    if (full) {
      f(LdsView<T, true>(lv[l]));
    } else {
      f(LdsView<T, false>(lv[l]));
    }
  };
  auto sweep = [&](int l, const T* u_, T* out_, T om) {
    const lds_cp<T> u = (lds_cp<T>)u_;
    const lds_p<T> out = (lds_p<T>)out_;
    with_view(l, [&](const auto& V) {
      for_cells(V.n, V.nx, V.ny, [&](int i, int x, int y, int z) {
        out[i] = u[i] + om * V.dinv(i) * (V.r[i] - V.lap(i, x, y, z, u));
      });
    });
    __syncthreads();
  };
  // FP32 levels resident in LDS with nx % 4 == 0: groups of 4 cells per
  // thread (the packed coefficients of a group in two 16-byte reads, the
  // neighbouring rows and planes in 16-byte reads), the operation order of
  // LdsView::lap per cell
  using V4 = vec_t<float, 4>;
  using U4 = vec_t<unsigned, 4>;
  using U2 = vec_t<unsigned, 2>;
  auto vec_ok = [&](int l) { return sizeof(T) == 4 && full && lv[l].nx % 4 == 0; };
  // f(g, i, x0, y, z) for the groups g = i / 4 of level l
  auto for_groups = [&](int l, auto f) {
    const int nx = lv[l].nx, ny = lv[l].ny, ng = (int)lv[l].n / 4;
    for (int g = tid; g < ng; g += nt) {
      const int i = 4 * g;
      const int x0 = i % nx, t = i / nx;
      f(g, i, x0, t % ny, t / ny);
    }
  };
  // (L u) of the 4 cells from i = 4 g (x0, y, z) of level l, u in LDS
  auto lap4 = [&](int l, const float* u_, int g, int i, int x0, int y, int z) -> V4 {
    const lds_cp<float> u = (lds_cp<float>)u_;
    const lds_cp<unsigned short> kc = (lds_cp<unsigned short>)(const unsigned short*)lv[l].kc;
    const int nx = lv[l].nx, ny = lv[l].ny, nz = lv[l].nz, pl = nx * ny, n = (int)lv[l].n;
    const U4 c0 = *(lds_cp<U4>)(kc + 16 * g), c1 = *(lds_cp<U4>)(kc + 16 * g + 8);
    const int ixm = x0 == 0 ? i + nx - 1 : i - 1, ixp = x0 + 4 == nx ? i - nx + 4 : i + 4;
    const int iym = y == 0 ? i + pl - nx : i - nx, iyp = y == ny - 1 ? i - pl + nx : i + nx;
    const int izm = z == 0 ? i + n - pl : i - pl, izp = z == nz - 1 ? i - n + pl : i + pl;
    const V4 w = *(lds_cp<V4>)(u + i);
    const V4 wym = *(lds_cp<V4>)(u + iym), wyp = *(lds_cp<V4>)(u + iyp);
    const V4 wzm = *(lds_cp<V4>)(u + izm), wzp = *(lds_cp<V4>)(u + izp);
    const float wxm = u[ixm], wxp = u[ixp];
    const U2 ym2 = *(lds_cp<U2>)(kc + 16 * (iym >> 2) + 4);
    const U2 zm2 = *(lds_cp<U2>)(kc + 16 * (izm >> 2) + 8);
    const float kxm = bflo(kc[16 * (ixm >> 2) + (ixm & 3)]);
    const V4 kx = bf4(c0.x, c0.y), ky = bf4(c0.z, c0.w), kz = bf4(c1.x, c1.y);
    const V4 kym = bf4(ym2.x, ym2.y), kzm = bf4(zm2.x, zm2.y);
    V4 s4;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const float up = v < 3 ? w[v + 1] : wxp, uw = v > 0 ? w[v - 1] : wxm;
      const float kw = v > 0 ? kx[v - 1] : kxm;
      s4[v] = kx[v] * (w[v] - up) + kw * (w[v] - uw) + ky[v] * (w[v] - wyp[v]) +
              kym[v] * (w[v] - wym[v]) + kz[v] * (w[v] - wzp[v]) + kzm[v] * (w[v] - wzm[v]);
    }
    if (lv[l].m) s4 += *(lds_cp<V4>)((const float*)lv[l].m + i) * w;
    return s4;
  };
  auto dinv4 = [&](int l, int g) {
    const U2 d = *(lds_cp<U2>)((lds_cp<unsigned short>)(const unsigned short*)lv[l].kc +
                               16 * g + 12);
    return bf4(d.x, d.y);
  };
  // out = u + om D^-1 (r - L u) on level l by groups, then a barrier
  auto sweep4 = [&](int l, const T* u, T* out, T om) {
    for_groups(l, [&](int g, int i, int x0, int y, int z) {
      const V4 uc = *(lds_cp<V4>)((const float*)u + i);
      const V4 rr = *(lds_cp<V4>)((const float*)lv[l].r + i);
      *(lds_p<V4>)((float*)out + i) =
          uc + (float)om * dinv4(l, g) * (rr - lap4(l, (const float*)u, g, i, x0, y, z));
    });
    __syncthreads();
  };
  for (int l = l0; l < lw && l < top; ++l) {
    const lds_p<T> zal = (lds_p<T>)za[l];
    const lds_p<T> zbl = (lds_p<T>)zb[l];
    if (vec_ok(l)) {
      for_groups(l, [&](int g, int i, int, int, int) {
        *(lds_p<V4>)((float*)zb[l] + i) =
            (float)wa * dinv4(l, g) * *(lds_cp<V4>)((const float*)lv[l].r + i);
      });
      __syncthreads();
      sweep4(l, zb[l], za[l], wb);
      for_groups(l, [&](int g, int i, int x0, int y, int z) {
        *(lds_p<V4>)((float*)zb[l] + i) = *(lds_cp<V4>)((const float*)lv[l].r + i) -
                                          lap4(l, (const float*)za[l], g, i, x0, y, z);
      });
      __syncthreads();
    } else {
    with_view(l, [&](const auto& V) {
      for (int i = tid; i < V.n; i += nt) zbl[i] = wa * V.dinv(i) * V.r[i];
    });
    __syncthreads();
    sweep(l, zb[l], za[l], wb);
    // restriction of r - L z into r of level l + 1: the residual of every
    // cell into the free second buffer, then the sums over the 2x2x2
    // children in a fixed order
    with_view(l, [&](const auto& V) {
      for_cells(V.n, V.nx, V.ny, [&](int i, int x, int y, int z) {
        zbl[i] = V.r[i] - V.lap(i, x, y, z, (lds_cp<T>)zal);
      });
    });
    __syncthreads();
    }
    const MgLevel& Lf = lv[l];
    with_view(l + 1, [&](const auto& C) {
      const int fnx = Lf.nx, fpl = Lf.nx * Lf.ny;
      for_cells(C.n, C.nx, C.ny, [&](int c, int X, int Y, int Z) {
        const int f0 = 2 * Z * fpl + 2 * Y * fnx + 2 * X;
        T sum = 0;
        for (int dz = 0; dz < 2; ++dz)
          for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) sum += zbl[f0 + dz * fpl + dy * fnx + dx];
        C.r[c] = sum;
      });
    });
    __syncthreads();
  }
  if (lw <= top) {
    if (tid < 64) {
      const int lane = tid;
      WaveLevel<T> A;
      A.load(lv[lw], lane);
      const T ra = lane < A.n ? ((const T*)lv[lw].r)[lane] : (T)0;
      T z;
      if (lw == top) {
        z = w * A.d * ra;
        for (int k = 1; k < sweeps; ++k) z = z + w * A.d * (ra - A.lap(z));
      } else {
        const MgLevel& Lf = lv[lw];
        const MgLevel& Lc = lv[lw + 1];
        z = wa * A.d * ra;
        z = z + wb * A.d * (ra - A.lap(z));
        const T res = ra - A.lap(z);
        WaveLevel<T> B;
        B.load(Lc, lane);
        const bool con = lane < B.n;
        const int C = con ? lane : 0;
        const int X = C % Lc.nx, Y = (C / Lc.nx) % Lc.ny, Z = C / (Lc.nx * Lc.ny);
        T rb = 0;
        for (int dz = 0; dz < 2; ++dz)
          for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
              const int f = ((2 * Z + dz) * Lf.ny + 2 * Y + dy) * Lf.nx + 2 * X + dx;
              rb += __shfl(res, f, 64);
            }
        if (!con) rb = (T)0;
        T zc = w * B.d * rb;
        for (int k = 1; k < sweeps; ++k) zc = zc + w * B.d * (rb - B.lap(zc));
        // prolongation of the coarsest correction, two sweeps
        const int c = lane < A.n ? lane : 0;
        const int x = c % Lf.nx, y = (c / Lf.nx) % Lf.ny, zz = c / (Lf.nx * Lf.ny);
        const T u =
            z + (T)kProlong * __shfl(zc, ((zz >> 1) * Lc.ny + (y >> 1)) * Lc.nx + (x >> 1), 64);
        z = u + wb * A.d * (ra - A.lap(u));
        z = z + wa * A.d * (ra - A.lap(z));
      }
      if (lane < A.n) za[lw][lane] = z;
    }
    __syncthreads();
  } else {
    // a coarsest level of more than 64 cells: block-wide sweeps
    const int n = (int)lv[top].n;
    T* a = za[top];
    T* b = zb[top];
    with_view(top, [&](const auto& V) {
      const lds_p<T> al = (lds_p<T>)a;
      for (int i = tid; i < n; i += nt) al[i] = w * V.dinv(i) * V.r[i];
    });
    __syncthreads();
    for (int k = 1; k < sweeps; ++k) {
      sweep(top, a, b, w);
      T* t = a;
      a = b;
      b = t;
    }
    if (a != za[top]) {
      for (int i = tid; i < n; i += nt) za[top][i] = a[i];
      __syncthreads();
    }
  }
  // the post-smoothed correction of level l ends in zb[l] (bit l of inb;
  // the levels solved above end in za)
  unsigned inb = 0;
  for (int l = (lw <= top ? lw : top) - 1; l >= l0; --l) {
    const MgLevel& Lf = lv[l];
    const MgLevel& Lc = lv[l + 1];
    const int n = (int)Lf.n;
    const lds_cp<T> zc = (lds_cp<T>)(((inb >> (l + 1)) & 1u) ? zb[l + 1] : za[l + 1]);
    const lds_cp<T> zf = (lds_cp<T>)za[l];
    const lds_p<T> u = (lds_p<T>)zb[l];
    for_cells(n, Lf.nx, Lf.ny, [&](int i, int x, int y, int z) {
      u[i] = zf[i] + (T)kProlong * zc[((z >> 1) * Lc.ny + (y >> 1)) * Lc.nx + (x >> 1)];
    });
    __syncthreads();
    if (vec_ok(l)) {
      sweep4(l, zb[l], za[l], wb);
      sweep4(l, za[l], zb[l], wa);
    } else {
      sweep(l, zb[l], za[l], wb);
      sweep(l, za[l], zb[l], wa);
    }
    inb |= 1u << l;
  }
  const T* __restrict__ res = ((inb >> l0) & 1u) ? zb[l0] : za[l0];
  T* out = (T*)s.lev[l0].z[l0 == top ? 0 : 1];  // global
  for (int i = tid; i < (int)lv[l0].n; i += nt) out[i] = res[i];
}

// (r, z) of the preconditioned CG: beta = (r,z) / (r,z)_old (0 after a start)
__global__ __launch_bounds__(kBlock) void k_fin_precond(const SysDev* __restrict__ sds,
                                                        int so) {
  const SysDev& s = sds[blockIdx.x];
  Scalars* sc = s.sc;
  if (s.nlev <= 0 || sc->done) return;
  if (so && !sc->first) return;
  double acc[1];
  reduce_partials<1, 0u>(acc, s.partials, s.nblk_pc);
  if (threadIdx.x == 0) {
    const double rz = acc[0];
    sc->beta = sc->first ? 0.0 : rz / sc->rho;
    sc->rho = rz;
    sc->rz = rz;
    sc->first = 0;
    if (!isfinite(rz)) {
      sc->status = 6;
      sc->done = 1;
    }
  }
}

// Level-0 coefficients of the V-cycle from the CG's conductances (a copy
// only where the levels store another type: the packed bfloat16 of FP32
// engines).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mg_level0(MgLevel L, const T* __restrict__ kx,
                                                      const T* __restrict__ ky,
                                                      const T* __restrict__ kz) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < L.n;
       i += (long long)gridDim.x * kBlock) {
    st_coef<T>(L, 0, i, kx[i]);
    st_coef<T>(L, 1, i, ky[i]);
    st_coef<T>(L, 2, i, kz[i]);
  }
}

// Coarse operator of level l+1 from level l (one thread per coarse cell):
// coarse face = scale * (sum of the 4 fine faces it covers), m = scale * sum
// of the 8 children, in T, stored in the levels' coefficient format.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mg_coarsen(MgLevel f, MgLevel c, T scale) {
  const MgCoefs<T> F(f);
  const long long n = c.n;
  const T* fm = (const T*)f.m;
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    const int x = (int)(i % c.nx);
    const long long t = i / c.nx;
    const int y = (int)(t % c.ny);
    const int z = (int)(t / c.ny);
    auto idx = [&](int fx, int fy, int fz) {
      return ((long long)fz * f.ny + fy) * f.nx + fx;
    };
    T sx = 0, sy = 0, sz = 0, sm = 0;
    for (int a = 0; a < 2; ++a)
      for (int bb = 0; bb < 2; ++bb) {
        sx += F.comp1(0, idx(2 * x + 1, 2 * y + a, 2 * z + bb));
        sy += F.comp1(1, idx(2 * x + a, 2 * y + 1, 2 * z + bb));
        sz += F.comp1(2, idx(2 * x + a, 2 * y + bb, 2 * z + 1));
      }
    if (fm) {
      for (int d = 0; d < 8; ++d) sm += fm[idx(2 * x + (d & 1), 2 * y + ((d >> 1) & 1), 2 * z + (d >> 2))];
      ((T*)c.m)[i] = scale * sm;
    }
    st_coef<T>(c, 0, i, scale * sx);
    st_coef<T>(c, 1, i, scale * sy);
    st_coef<T>(c, 2, i, scale * sz);
  }
}

// D^-1 of a level from its stored coefficients (the diagonal of the operator
// the sweeps apply).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mg_dinv(MgLevel L) {
  const MgCoefs<T> co(L);
  const T* m = (const T*)L.m;
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < L.n;
       i += (long long)gridDim.x * kBlock) {
    const int x = (int)(i % L.nx);
    const long long t = i / L.nx;
    const int y = (int)(t % L.ny);
    const int z = (int)(t / L.ny);
    const long long pl = (long long)L.nx * L.ny;
    T d = co.comp1(0, i) + co.comp1(0, (long long)z * pl + (long long)y * L.nx + wrapm(x, L.nx)) +
          co.comp1(1, i) + co.comp1(1, (long long)z * pl + (long long)wrapm(y, L.ny) * L.nx + x) +
          co.comp1(2, i) + co.comp1(2, (long long)wrapm(z, L.nz) * pl + (long long)y * L.nx + x);
    if (m) d += m[i];
    st_coef<T>(L, 3, i, d != (T)0 ? (T)1 / d : (T)0);
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
  const int sh = s.f7sh;  // 0: SoA arrays, 3: the Expr's AoS (8 doubles per cell)
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
      double v = u[c] * s.f7[0][c << sh] + s.f7[7][c << sh];
      v += u[base + row + xw] * s.f7[1][c << sh];
      v += u[base + row + xe] * s.f7[2][c << sh];
      v += u[base + rowm + xx] * s.f7[3][c << sh];
      v += u[base + rowp + xx] * s.f7[4][c << sh];
      v += u[(long long)wrapm(z, nz) * pl + row + xx] * s.f7[5][c << sh];
      v += u[(long long)wrapp(z, nz) * pl + row + xx] * s.f7[6][c << sh];
      const double rv = -v;
      s.rr[c] = rv;
      s.pp[0][c] = rv;
      acc[0] += rv * rv;
    }
  }
  block_partial<1, 0u>(acc, s.partials, rc.nblk, lb);
}

__global__ __launch_bounds__(kBlock) void k_fin_ref_init(const SysDev* __restrict__ sds) {
  const SysDev& s = sds[blockIdx.x];
  Scalars* sc = s.sc;
  if (sc->done) return;
  double acc[1];
  reduce_partials<1, 0u>(acc, s.partials, s.nblk_st);
  if (threadIdx.x == 0) {
    sc->rho = acc[0];
    sc->cur = 0;
    sc->first = 1;
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
  const int sh = s.f7sh;  // 0: SoA arrays, 3: the Expr's AoS (8 doubles per cell)
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
      double v = pc * s.f7[0][c << sh];
      v += pn(base + row + xw) * s.f7[1][c << sh];
      v += pn(base + row + xe) * s.f7[2][c << sh];
      v += pn(base + rowm + xx) * s.f7[3][c << sh];
      v += pn(base + rowp + xx) * s.f7[4][c << sh];
      v += pn((long long)wrapm(z, nz) * pl + row + xx) * s.f7[5][c << sh];
      v += pn((long long)wrapp(z, nz) * pl + row + xx) * s.f7[6][c << sh];
      pout[c] = pc;
      s.qq[c] = v;
      acc[0] += pc * v;
    }
  }
  block_partial<1, 0u>(acc, s.partials, rc.nblk, lb);
}

__global__ __launch_bounds__(kBlock) void k_fin_ref_spmv(const SysDev* __restrict__ sds) {
  const SysDev& s = sds[blockIdx.x];
  Scalars* sc = s.sc;
  if (sc->done) return;
  double acc[1];
  reduce_partials<1, 0u>(acc, s.partials, s.nblk_st);
  if (threadIdx.x == 0) {
    sc->alpha = sc->rho / (acc[0] + 1e-100);
    sc->cur ^= 1;
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
  block_partial<2, 2u>(acc, s.partials, rc.nblk, lb);
}

__global__ __launch_bounds__(kBlock) void k_fin_ref_update(const SysDev* __restrict__ sds) {
  const SysDev& s = sds[blockIdx.x];
  Scalars* sc = s.sc;
  if (sc->done) return;
  double acc[2];
  reduce_partials<2, 2u>(acc, s.partials, s.nblk_sm);
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
  const int sh = s.f7sh;  // 0: SoA arrays, 3: the Expr's AoS (8 doubles per cell)
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
      double nd = s.f7[7][c << sh];
      nd += u[base + row + xw] * s.f7[1][c << sh];
      nd += u[base + row + xe] * s.f7[2][c << sh];
      nd += u[base + rowm + xx] * s.f7[3][c << sh];
      nd += u[base + rowp + xx] * s.f7[4][c << sh];
      nd += u[(long long)wrapm(z, nz) * pl + row + xx] * s.f7[5][c << sh];
      nd += u[(long long)wrapp(z, nz) * pl + row + xx] * s.f7[6][c << sh];
      const double v = -nd / s.f7[0][c << sh];
      un[c] = v;
      acc[0] = fmax(acc[0], fabs(v - u[c]));
    }
  }
  block_partial<1, 1u>(acc, s.partials, rc.nblk, lb);
}

__global__ __launch_bounds__(kBlock) void k_fin_jacobi(const SysDev* __restrict__ sds) {
  const SysDev& s = sds[blockIdx.x];
  Scalars* sc = s.sc;
  if (sc->done) return;
  double acc[1];
  reduce_partials<1, 1u>(acc, s.partials, s.nblk_st);
  if (threadIdx.x == 0) {
    const double res = acc[0];
    sc->residual = res;
    sc->iter += 1;
    sc->cur ^= 1;
    if (sc->iter >= sc->miniter && (sc->iter > sc->maxiter || res < sc->tol)) {
      sc->converged = res < sc->tol ? 1 : 0;
      sc->done = 1;
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
  const int sh = s.f7sh;  // 0: SoA arrays, 3: the Expr's AoS (8 doubles per cell)
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
      double v = u[c] * s.f7[0][c << sh] + s.f7[7][c << sh];
      v += u[base + row + xw] * s.f7[1][c << sh];
      v += u[base + row + xe] * s.f7[2][c << sh];
      v += u[base + rowm + xx] * s.f7[3][c << sh];
      v += u[base + rowp + xx] * s.f7[4][c << sh];
      v += u[(long long)wrapm(z, nz) * pl + row + xx] * s.f7[5][c << sh];
      v += u[(long long)wrapp(z, nz) * pl + row + xx] * s.f7[6][c << sh];
      acc[0] += v * v;
      acc[1] += s.f7[7][c << sh] * s.f7[7][c << sh];
      acc[2] = fmax(acc[2], fabs(v));
    }
  }
  block_partial<3, 4u>(acc, s.partials, rc.nblk, lb);
}

__global__ __launch_bounds__(kBlock) void k_fin_true_res_f7(const SysDev* __restrict__ sds) {
  const SysDev& s = sds[blockIdx.x];
  Scalars* sc = s.sc;
  double acc[3];
  reduce_partials<3, 4u>(acc, s.partials, s.nblk_ru);
  if (threadIdx.x == 0) {
    sc->relres = acc[1] > 0 ? sqrt(acc[0] / acc[1]) : sqrt(acc[0]);
    sc->res_l2 = sqrt(acc[0] / sc->h3);
    sc->res_max = acc[2] / sc->h3;
  }
}

#pragma clang fp contract(on)

// ===========================================================================
// System conversion (set_system): AoS Expr with strides -> SoA layouts.
// ===========================================================================

// flags: 1 non-finite, 2 asymmetric, 4 zero diagonal, 16 positive diagonal
// seen, 32 negative diagonal seen, 64 non-zero row sum (mass term), 128 e0
// other than -(((((e1 + e2) + e3) + e4) + e5) + e6) bit for bit
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
    if (e[0] != -(((((e[1] + e[2]) + e[3]) + e[4]) + e[5]) + e[6])) f |= 128u;
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
    if (S) S[i] = (T)(1.0 / sqrt(fabs(e[0])));
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
  (void)counter;
  block_partial<1, 0u>(acc, partials, gridDim.x, blockIdx.x);
}

__global__ __launch_bounds__(kBlock) void k_fin_sumsq(const double* __restrict__ partials,
                                                      int nblk, double* __restrict__ out) {
  double acc[1];
  reduce_partials<1, 0u>(acc, partials, nblk);
  if (threadIdx.x == 0) *out = acc[0];
}

// Correlated right-hand sides for served request streams:
//   b <- b + amp * max|b| * (eta - mean(eta)),  eta_i = 2 u(key, i) - 1
// with u the splitmix64 counter generator (as in the synthetic systems).
__device__ __forceinline__ unsigned long long d_sm64(unsigned long long x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__device__ __forceinline__ double d_eta(unsigned long long key, long long i) {
  return 2.0 * (double)(d_sm64(key + (unsigned long long)i) >> 11) * 0x1.0p-53 - 1.0;
}

// partials: sum eta, max |b|
__global__ __launch_bounds__(kBlock) void k_perturb_stats(const double* __restrict__ b,
                                                          long long n, unsigned long long key,
                                                          double* __restrict__ partials) {
  double acc[2] = {0.0, 0.0};
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    acc[0] += d_eta(key, i);
    acc[1] = fmax(acc[1], fabs(b[i]));
  }
  block_partial<2, 2u>(acc, partials, gridDim.x, blockIdx.x);
}

__global__ __launch_bounds__(kBlock) void k_fin_perturb_stats(double* __restrict__ partials,
                                                              int nblk, long long n) {
  double acc[2];
  reduce_partials<2, 2u>(acc, partials, nblk);
  if (threadIdx.x == 0) {
    partials[2 * nblk] = acc[0] / (double)n;  // mean eta
    partials[2 * nblk + 1] = acc[1];          // max |b|
  }
}

__global__ __launch_bounds__(kBlock) void k_perturb_apply(double* __restrict__ b, long long n,
                                                          unsigned long long key, double amp,
                                                          const double* __restrict__ stats) {
  const double mean = stats[0], bmax = stats[1];
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    b[i] += amp * bmax * (d_eta(key, i) - mean);
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill(T* __restrict__ p, long long n, T v) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    p[i] = v;
  }
}

// block -> record map of a one-system launch set with three records:
// n0 workgroups of record 0, then n1 of record 1, then n2 of record 2
__global__ __launch_bounds__(kBlock) void k_b2r3(int* __restrict__ b2r, int n0, int n1,
                                                 int n2) {
  const int n = n0 + n1 + n2;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    b2r[i] = i < n0 ? 0 : (i < n0 + n1 ? 1 : 2);
}

// x -= sub in place (a warm start folding a pending mean into the solution)
__global__ __launch_bounds__(kBlock) void k_shift(double* __restrict__ x, long long n,
                                                  double sub) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    x[i] -= sub;
  }
}

// Difference of a solution from a reference solution, d = (x - sub) - xref:
// partials of sum d (pass 0), then of max |d - mean| with the mean read from
// stats[0] (pass 1; stats NULL: mean 0).
// max |d - mean d| over the cells, d = (x - sub) - xref (mean 0 unless
// remove_mean), in one pass: per workgroup the sum, the max and the max of -d.
__global__ __launch_bounds__(kBlock) void k_diff(const double* __restrict__ x,
                                                 const double* __restrict__ xref, long long n,
                                                 double sub, double* __restrict__ partials) {
  double acc[3] = {0.0, -INFINITY, -INFINITY};
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    const double d = (x[i] - sub) - xref[i];
    acc[0] += d;
    acc[1] = fmax(acc[1], d);
    acc[2] = fmax(acc[2], -d);
  }
  block_partial<3, 6u>(acc, partials, gridDim.x, blockIdx.x);
}

// out[0] = mean d (0 unless remove_mean), out[1] = max(max d - mean, mean -
// min d, 0): the largest |d - mean| bit for bit, since rounding d - mean is
// monotonic in d and symmetric in sign.
__global__ __launch_bounds__(kBlock) void k_fin_diff(const double* __restrict__ partials,
                                                     int nblk, int remove_mean, double scale,
                                                     double* __restrict__ out) {
  double sum[1];
  reduce_partials<1, 0u>(sum, partials, nblk);
  double mx[2] = {-INFINITY, -INFINITY};
  for (int i = threadIdx.x; i < nblk; i += kBlock) {
    mx[0] = fmax(mx[0], partials[nblk + i]);
    mx[1] = fmax(mx[1], partials[2 * nblk + i]);
  }
  block_reduce<2, 3u>(mx);
  if (threadIdx.x == 0) {
    const double mean = remove_mean ? sum[0] * scale : 0.0;
    out[0] = mean;
    out[1] = fmax(0.0, fmax(mx[0] - mean, mean - (-mx[1])));
  }
}

// ===========================================================================
// launchers
// ===========================================================================

#define APH_LAUNCH(kern, grid, stream, ...) \
  hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), 0, stream, __VA_ARGS__)

int iter_kernels(int method, int pc) {
  return method == kConjugate ? 4 : (method == kJacobi ? 2 : (pc ? 9 : 8));
}

int graph_kernels(int method, int pc) { return method == kPcg && !pc ? 4 : iter_kernels(method, pc); }

// One V-cycle over every system of the launch (levels of systems with fewer
// levels drop out of the deeper launches), then (r, z) and beta.
template <typename T, int MODE, bool DOT, bool L0>
static void launch_sweep_l(const LaunchSet& L, hipStream_t st, int l, int zin, int zout,
                           int wl, int addzs) {
  switch (L.vx_l[l]) {
    case 4:
      APH_LAUNCH((k_mg_sweep<T, 4, MODE, DOT, L0>), L.nlf[l], st, L.sds, L.rec_lf[l],
                 L.b2r_lf[l], l, zin, zout, wl, L.start_only, addzs);
      break;
    case 2:
      APH_LAUNCH((k_mg_sweep<T, 2, MODE, DOT, L0>), L.nlf[l], st, L.sds, L.rec_lf[l],
                 L.b2r_lf[l], l, zin, zout, wl, L.start_only, addzs);
      break;
    default:
      if constexpr (MODE == 3) {
        APH_LAUNCH((k_mg_restrict<T, L0>), L.nlc[l], st, L.sds, L.rec_lc[l], L.b2r_lc[l], l,
                   zin, wl, L.start_only);
      } else {
        APH_LAUNCH((k_mg_sweep<T, 1, MODE, DOT, L0>), L.nlf[l], st, L.sds, L.rec_lf[l],
                   L.b2r_lf[l], l, zin, zout, wl, L.start_only, addzs);
      }
      break;
  }
}

// pre-smoothing + restriction of level l, fused, for the systems in rec_lp[l]
template <typename T, bool L0>
static void launch_pre_l(const LaunchSet& L, hipStream_t st, int l, int wl) {
  const dim3 g(L.nlp[l]), t(kPostBlock);
  if (L.vx_p[l] == 4) {
    if constexpr (sizeof(T) == 4) {
      hipLaunchKernelGGL((k_mg_pre4<L0>), g, t, 0, st, L.sds, L.rec_lp[l], L.b2r_lp[l], l, wl,
                         L.start_only);
    }
  } else {
    hipLaunchKernelGGL((k_mg_pre<T, 2, L0>), g, t, 0, st, L.sds, L.rec_lp[l], L.b2r_lp[l], l, wl,
                       L.start_only);
  }
}

// both post-smoothing sweeps of level l, fused, for the systems in rec_lp[l]
template <typename T, bool DOT, bool L0>
static void launch_post_l(const LaunchSet& L, hipStream_t st, int l, int wl, int addzs) {
  const dim3 g(L.nlp[l]), t(kPostBlock);
  // FP32 engines use 4 cells per thread, FP64 engines 2 (the LDS planes of a
  // 1024-thread workgroup: 5 x 1024 x VX x sizeof(T) bytes)
  switch (L.vx_p[l]) {
    case 4:
      if constexpr (sizeof(T) == 4) {
        hipLaunchKernelGGL((k_mg_post4<DOT, L0>), g, t, 0, st, L.sds, L.rec_lp[l],
                           L.b2r_lp[l], l, wl, L.start_only, addzs);
      }
      break;
    case 2:
      hipLaunchKernelGGL((k_mg_post<T, 2, DOT, L0>), g, t, 0, st, L.sds, L.rec_lp[l], L.b2r_lp[l],
                         l, wl, L.start_only, addzs);
      break;
    default:
      hipLaunchKernelGGL((k_mg_post<T, 1, DOT, L0>), g, t, 0, st, L.sds, L.rec_lp[l], L.b2r_lp[l],
                         l, wl, L.start_only, addzs);
      break;
  }
}

template <typename T, int MODE, bool DOT>
static void launch_sweep(const LaunchSet& L, hipStream_t st, int l, int zin, int zout, int wl,
                         int addzs = 0) {
  if (l == 0) launch_sweep_l<T, MODE, DOT, true>(L, st, l, zin, zout, wl, addzs);
  else launch_sweep_l<T, MODE, DOT, false>(L, st, l, zin, zout, wl, addzs);
}

// The cycle from level l down, recursively.  Per level above the coarsest:
// the zero-guess sweep fused with the second pre-smoothing sweep (-> z[1]),
// restriction of r - L z[1], the coarse correction (twice on W levels: the
// first saved to zs by the in-place residual r -= L z, MODE 4, then a second
// cycle), the prolongation of the correction(s) fused with the first
// post-smoothing sweep (z[1] -> z[0]) and one more sweep (z[0] -> z[1]; on
// level 0 it also forms (r, z)).  The correction of every level ends in z[1],
// of the coarsest in z[0].  wl: the innermost W level whose second correction
// this is (-1: none); each system takes part only while wl < its mg_w.
// second: this cycle is the second correction of level l - 1; its last sweep
// adds the first correction (zs of level l), so z[1] of level l holds the sum
// the prolongation needs.  (A second correction starts at a level <= mg_w,
// above the bottom kernel's first level, and below the coarsest.)
// `mark(kind)` precedes every launch (profiling): 0 = level-0 pre-smoothing
// sweep, 1 = level-0 restriction, 2 = coarser-level launch, 3 = bottom
// levels (k_mg_bottom), 5 and 6 = level-0 first and second post-smoothing
// sweeps (5 alone: both, fused).
template <typename T, typename Mark>
static void launch_cycle(const LaunchSet& L, hipStream_t st, int l, int wl, bool second,
                         Mark& mark) {
  if (L.bottom_at[l]) {
    mark(3);
    hipLaunchKernelGGL(k_mg_bottom<T>, dim3(L.ns), dim3(kBottomThreads), L.coarse_lds, st,
                       L.sds, L.coarse_sweeps, l, wl, L.start_only);
  }
  if (l >= L.deep) return;  // every system's cycle ends at or above level l
  // the pre-smoothed iterate (zout / zin -1): z[0] of a system whose level
  // fuses the post-smoothing (which then writes z[1]), else z[1] (the first
  // post-sweep writes z[0], the second z[1])
  if (L.post_fused[l]) {
    mark(l == 0 ? 0 : 2);
    if (l == 0) launch_pre_l<T, true>(L, st, l, wl);
    else launch_pre_l<T, false>(L, st, l, wl);
  }
  if (L.post_plain[l]) {
    mark(l == 0 ? 0 : 2);
    launch_sweep<T, 1, false>(L, st, l, 0, -1, wl);
    mark(l == 0 ? 1 : 2);
    launch_sweep<T, 3, false>(L, st, l, -1, -1, wl);
  }
  if (L.post_tile[l]) {
    mark(2);
    if constexpr (sizeof(T) == 4) {
      hipLaunchKernelGGL(k_mg_tile<true>, dim3(L.nlt[l]), dim3(kTileBlock), 0, st, L.sds,
                         L.rec_lt[l], L.b2r_lt[l], l, wl, L.start_only, 0);
    }
  }
  if (l == 0 && L.below0) {
    // event profiling: the coarser levels as one replay (one event)
    mark(2);
    (void)hipGraphLaunch(L.below0, st);
  } else {
    launch_cycle<T>(L, st, l + 1, wl, false, mark);
    if (l < L.mg_w) {
      mark(2);
      launch_sweep<T, 4, false>(L, st, l + 1, 1, 1, l);
      launch_cycle<T>(L, st, l + 1, l, true, mark);
    }
  }
  if (L.post_fused[l]) {
    mark(l == 0 ? 5 : 2);
    if (l == 0) launch_post_l<T, true, true>(L, st, l, wl, 0);
    else launch_post_l<T, false, false>(L, st, l, wl, second ? 1 : 0);
  }
  if (L.post_tile[l]) {
    mark(2);
    if constexpr (sizeof(T) == 4) {
      hipLaunchKernelGGL(k_mg_tile<false>, dim3(L.nlt[l]), dim3(kTileBlock), 0, st, L.sds,
                         L.rec_lt[l], L.b2r_lt[l], l, wl, L.start_only, second ? 1 : 0);
    }
  }
  if (L.post_plain[l]) {
    mark(l == 0 ? 5 : 2);
    launch_sweep<T, 2, false>(L, st, l, 1, 0, wl);
    mark(l == 0 ? 6 : 2);
    if (l == 0) launch_sweep<T, 0, true>(L, st, l, 0, 1, wl);
    else launch_sweep<T, 0, false>(L, st, l, 0, 1, wl, second ? 1 : 0);
  }
}

// The preconditioner z = M r on level 0 and the (r, z) finalize; `ev`
// (profiling) records an event before every launch and after the last,
// `kinds` receives the kinds of launch_cycle and 4 = finalize.
template <typename T>
static int launch_vcycle(const LaunchSet& L, hipStream_t st, hipEvent_t* ev = nullptr,
                         int* kinds = nullptr) {
  int k = 0;
  auto mark = [&](int kind) {
    if (ev) {
      (void)hipEventRecord(ev[k], st);
      kinds[k] = kind;
    }
    ++k;
  };
  launch_cycle<T>(L, st, 0, -1, false, mark);
  mark(4);
  APH_LAUNCH(k_fin_precond, L.ns, st, L.sds, L.start_only);
  if (ev) (void)hipEventRecord(ev[k], st);
  return k;
}

hipError_t launch_vcycle_profiled(int fp32, const LaunchSet& L, hipStream_t st, hipEvent_t* ev,
                                  int* kinds, int* nlaunch) {
  *nlaunch = fp32 ? launch_vcycle<float>(L, st, ev, kinds)
                  : launch_vcycle<double>(L, st, ev, kinds);
  return hipGetLastError();
}

hipError_t capture_below0(int fp32, const LaunchSet& L, hipStream_t st) {
  LaunchSet Lb = L;
  Lb.below0 = nullptr;
  auto none = [](int) {};
  if (L.deep >= 1) {
    if (fp32) {
      launch_cycle<float>(Lb, st, 1, -1, false, none);
      if (0 < Lb.mg_w) {
        launch_sweep<float, 4, false>(Lb, st, 1, 1, 1, 0);
        launch_cycle<float>(Lb, st, 1, 0, true, none);
      }
    } else {
      launch_cycle<double>(Lb, st, 1, -1, false, none);
      if (0 < Lb.mg_w) {
        launch_sweep<double, 4, false>(Lb, st, 1, 1, 1, 0);
        launch_cycle<double>(Lb, st, 1, 0, true, none);
      }
    }
  }
  return hipGetLastError();
}

static int cycle_launches(const LaunchSet& L, int l) {
  int n = L.bottom_at[l] ? 1 : 0;
  if (l >= L.deep) return n;
  n += (L.post_fused[l] ? 2 : 0) + (L.post_plain[l] ? 4 : 0) + (L.post_tile[l] ? 2 : 0) +
       cycle_launches(L, l + 1);
  if (l < L.mg_w) n += 1 + cycle_launches(L, l + 1);
  return n;
}

int vcycle_launches(const LaunchSet& L) { return cycle_launches(L, 0) + 1; }

template <typename T, int VX>
static void launch_pcg_kernel(const LaunchSet& L, hipStream_t st, int j) {
  switch (j) {
    case 0:
      if (L.pc) APH_LAUNCH((k_pcg_spmv<T, VX, false, true>), L.nst, st, L.sds, L.rec_st, L.b2r_st);
      else APH_LAUNCH((k_pcg_spmv<T, VX, true, true>), L.nst, st, L.sds, L.rec_st, L.b2r_st);
      break;
    case 8: launch_vcycle<T>(L, st); break;
    case 1: APH_LAUNCH(k_fin_pcg_spmv, L.ns, st, L.sds); break;
    case 2:
      APH_LAUNCH((k_pcg_update<T, VX>), L.nsm, st, L.sds, L.rec_sm, L.b2r_sm);
      break;
    case 3: APH_LAUNCH(k_fin_pcg_update, L.ns, st, L.sds); break;
    case 4:
      APH_LAUNCH((k_ru_accum<T, VX>), L.nsm, st, L.sds, L.rec_sm, L.b2r_sm);
      break;
    case 5: APH_LAUNCH(k_fin_ru_accum, L.ns, st, L.sds); break;
    case 6:
      if (L.vx_ru == 2) APH_LAUNCH(k_ru_resid2<T>, L.nru, st, L.sds, L.rec_ru, L.b2r_ru);
      else APH_LAUNCH((k_ru_resid<T, 1>), L.nru, st, L.sds, L.rec_ru, L.b2r_ru);
      break;
    default: APH_LAUNCH(k_fin_ru_resid, L.ns, st, L.sds); break;
  }
}

hipError_t launch_kernel(int method, int fp32, int vx, const LaunchSet& L, hipStream_t st,
                         int j) {
  if (method == kConjugate) {
    switch (j) {
      case 0: APH_LAUNCH(k_ref_spmv, L.nst, st, L.sds, L.rec_st, L.b2r_st); break;
      case 1: APH_LAUNCH(k_fin_ref_spmv, L.ns, st, L.sds); break;
      case 2: APH_LAUNCH(k_ref_update, L.nsm, st, L.sds, L.rec_sm, L.b2r_sm); break;
      default: APH_LAUNCH(k_fin_ref_update, L.ns, st, L.sds); break;
    }
  } else if (method == kJacobi) {
    if (j == 0) APH_LAUNCH(k_jacobi, L.nst, st, L.sds, L.rec_st, L.b2r_st);
    else APH_LAUNCH(k_fin_jacobi, L.ns, st, L.sds);
  } else if (fp32) {
    if (vx == 4) launch_pcg_kernel<float, 4>(L, st, j);
    else if (vx == 2) launch_pcg_kernel<float, 2>(L, st, j);
    else launch_pcg_kernel<float, 1>(L, st, j);
  } else {
    if (vx == 2) launch_pcg_kernel<double, 2>(L, st, j);
    else launch_pcg_kernel<double, 1>(L, st, j);
  }
  return hipGetLastError();
}

hipError_t launch_iter(int method, int fp32, int vx, const LaunchSet& L, hipStream_t st) {
  for (int j = 0; j < iter_kernels(method, L.pc); ++j) {
    const hipError_t e = launch_kernel(method, fp32, vx, L, st, j);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_start(int method, int fp32, int vx, const LaunchSet& L, hipStream_t st) {
  if (method == kConjugate) {
    APH_LAUNCH(k_ref_init, L.nst, st, L.sds, L.rec_st, L.b2r_st);
    APH_LAUNCH(k_fin_ref_init, L.ns, st, L.sds);
  } else if (method == kPcg) {
    // the start is a reliable update (ru = 1): x64 = x0, r^ = S (b - A x0),
    // followed by z = M^-1 r^ with the multigrid preconditioner.  Systems
    // already iterating (continuous batching) are at an iteration boundary and
    // take no part: the reliable-update kernels skip them (ru = 0), the cycle
    // and its finalize are launched start-only (first == 1).
    LaunchSet Ls = L;
    Ls.start_only = 1;
    Ls.below0 = nullptr;
    for (int j = 4; j < iter_kernels(method, L.pc); ++j) {
      const hipError_t e = launch_kernel(method, fp32, vx, Ls, st, j);
      if (e != hipSuccess) return e;
    }
  }
  return hipGetLastError();
}

static int grid_for(long long n);

template <typename T>
static hipError_t mg_setup(const MgLevel* lv, int nlev, const void* kx, const void* ky,
                           const void* kz, hipStream_t st) {
  if (std::is_same<T, float>::value) {
    APH_LAUNCH(k_mg_level0<T>, grid_for(lv[0].n), st, lv[0], (const T*)kx, (const T*)ky,
               (const T*)kz);
  }
  for (int l = 0; l < nlev; ++l) {
    if (l > 0) {
      APH_LAUNCH(k_mg_coarsen<T>, grid_for(lv[l].n), st, lv[l - 1], lv[l], (T)0.5);
    }
    APH_LAUNCH(k_mg_dinv<T>, grid_for(lv[l].n), st, lv[l]);
  }
  return hipGetLastError();
}

hipError_t launch_mg_setup(int fp32, const MgLevel* levels, int nlev, const void* kx,
                           const void* ky, const void* kz, hipStream_t st) {
  return fp32 ? mg_setup<float>(levels, nlev, kx, ky, kz, st)
              : mg_setup<double>(levels, nlev, kx, ky, kz, st);
}

hipError_t launch_true_res_f7(const LaunchSet& L, hipStream_t st) {
  APH_LAUNCH(k_true_res_f7, L.nru, st, L.sds, L.rec_ru, L.b2r_ru);
  APH_LAUNCH(k_fin_true_res_f7, L.ns, st, L.sds);
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
  APH_LAUNCH(k_fin_sumsq, 1, st, partials, g, out);
  return hipGetLastError();
}

hipError_t launch_b2r3(int* b2r, int n0, int n1, int n2, hipStream_t st) {
  APH_LAUNCH(k_b2r3, grid_for((long long)n0 + n1 + n2), st, b2r, n0, n1, n2);
  return hipGetLastError();
}

hipError_t launch_shift(double* x, long long n, double sub, hipStream_t st) {
  APH_LAUNCH(k_shift, grid_for(n), st, x, n, sub);
  return hipGetLastError();
}

hipError_t launch_max_diff(const double* x, const double* xref, long long n, double sub,
                           int remove_mean, double* partials, double* out, hipStream_t st) {
  int g = grid_for(n);
  if (g > 1024) g = 1024;  // partials: 3 x 1024 doubles
  APH_LAUNCH(k_diff, g, st, x, xref, n, sub, partials);
  APH_LAUNCH(k_fin_diff, 1, st, partials, g, remove_mean, 1.0 / (double)n, out);
  return hipGetLastError();
}

// ===========================================================================
// Projection around the solve (src/solver/proj.ipp:343-403): the pressure
// system from density, face volume fluxes and source, and the corrected face
// fluxes for the solution.  FP64 in the reference's operation order, no FMA
// contraction, so the assembled Expr is bit-identical to the host assembly.
// Face fields: 3 arrays of n values (x, y, z); entry c of direction d is the
// lower face of cell c.  The face denominator K_f of s = -area / K_f * dt is
// the harmonic face density from rho (Proj) or given per face in fk (SIMPLE's
// momentum coefficient, src/solver/simple.ipp:214-226); one of rho and fk is
// null.
// ===========================================================================
__device__ __forceinline__ double face_den(const double* __restrict__ rho,
                                           const double* __restrict__ fk, long long n, int d,
                                           long long cm, long long cp) {
#pragma clang fp contract(off)
  if (fk) return fk[(long long)d * n + cp];
  const double inv = (1.0 / rho[cp] + 1.0 / rho[cm]) * 0.5;
  return 1.0 / inv;
}

__global__ __launch_bounds__(kBlock) void k_assemble(
    const double* __restrict__ rho, const double* __restrict__ fk, const double* __restrict__ ffv,
    const double* __restrict__ src, int nx, int ny, int nz, double cell, double dt, int sign,
    double* __restrict__ expr) {
#pragma clang fp contract(off)
  const long long n = (long long)nx * ny * nz;
  const double h = cell;
  const double vol = h * h * h;
  const double area = vol / h;
  const double a = 1.0 / h;
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    const int x = (int)(i % nx);
    const long long t = i / nx;
    const int y = (int)(t % ny);
    const int z = (int)(t / ny);
    const long long pl = (long long)nx * ny;
    const long long nb[6] = {i - x + wrapm(x, nx), i - x + wrapp(x, nx),
                             i + (long long)(wrapm(y, ny) - y) * nx,
                             i + (long long)(wrapp(y, ny) - y) * nx,
                             i + (long long)(wrapm(z, nz) - z) * pl,
                             i + (long long)(wrapp(z, nz) - z) * pl};
    double e[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const bool plus = q & 1;
      const int d = q >> 1;
      const long long cm = plus ? i : nb[q];
      const long long cp = plus ? nb[q] : i;
      const double dens = face_den(rho, fk, n, d, cm, cp);
      const double s = -area / dens * dt;
      double v0 = -a * s, v1 = a * s, v2 = 0.0 * s;
      v2 += ffv[(long long)d * n + cp];
      if (!plus) {
        v0 = -v0;
        v1 = -v1;
        v2 = -v2;
      }
      e[0] += plus ? v0 : v1;
      e[1 + q] += plus ? v1 : v0;
      e[7] += v2;
    }
    if (src) e[7] -= src[i] * vol;
    double* out = expr + 8 * i;
#pragma unroll
    for (int k = 0; k < 8; ++k) out[k] = sign > 0 ? e[k] : -e[k];
  }
}

// out[f] = p[cm] * ffe[0] + p[cp] * ffe[1] + ffe[2] (UEB::Eval of an
// ExprFace, src/solver/approx_eb.h:338-344) with ffe of the assembly.
__global__ __launch_bounds__(kBlock) void k_flux_correct(
    const double* __restrict__ p, const double* __restrict__ rho, const double* __restrict__ fk,
    const double* __restrict__ ffv, double* __restrict__ out, int nx, int ny, int nz,
    double cell, double dt) {
#pragma clang fp contract(off)
  const long long n = (long long)nx * ny * nz;
  const double h = cell;
  const double vol = h * h * h;
  const double area = vol / h;
  const double a = 1.0 / h;
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    const int x = (int)(i % nx);
    const long long t = i / nx;
    const int y = (int)(t % ny);
    const int z = (int)(t / ny);
    const long long pl = (long long)nx * ny;
    const long long lower[3] = {i - x + wrapm(x, nx), i + (long long)(wrapm(y, ny) - y) * nx,
                                i + (long long)(wrapm(z, nz) - z) * pl};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const long long cm = lower[d];
      const double dens = face_den(rho, fk, n, d, cm, i);
      const double s = -area / dens * dt;
      const double e0 = -a * s, e1 = a * s;
      double e2 = 0.0 * s;
      e2 += ffv[(long long)d * n + i];
      const double t0 = p[cm] * e0;
      const double t1 = p[i] * e1;
      out[(long long)d * n + i] = t0 + t1 + e2;
    }
  }
}

// Plain copy (the practical HBM bandwidth a streaming kernel reaches): each
// lane moves 4 x 16 bytes per step (four loads in flight), grid-stride.
__global__ __launch_bounds__(kBlock) void k_copy(const float4* __restrict__ src,
                                                 float4* __restrict__ dst, long long n) {
  const long long stride = (long long)gridDim.x * kBlock;
  long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
    const float4 a = src[i], b = src[i + stride], c = src[i + 2 * stride],
                 d = src[i + 3 * stride];
    dst[i] = a;
    dst[i + stride] = b;
    dst[i + 2 * stride] = c;
    dst[i + 3 * stride] = d;
  }
  for (; i < n; i += stride) dst[i] = src[i];
}

hipError_t launch_copy(const void* src, void* dst, long long n16, hipStream_t st) {
  if (n16 <= 0) return hipSuccess;
  long long g = (n16 + kBlock - 1) / kBlock;
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(k_copy, dim3((unsigned)g), dim3(kBlock), 0, st, (const float4*)src,
                     (float4*)dst, n16);
  return hipGetLastError();
}

template <typename Ti, typename To>
__global__ __launch_bounds__(kBlock) void k_cast(const Ti* __restrict__ src, To* __restrict__ dst,
                                                 long long n) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    dst[i] = (To)src[i];
  }
}

hipError_t launch_cast(const void* src, int src_bytes, void* dst, int dst_bytes, long long n,
                       hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (src_bytes == 8 && dst_bytes == 4) {
    APH_LAUNCH((k_cast<double, float>), grid_for(n), st, (const double*)src, (float*)dst, n);
  } else if (src_bytes == 4 && dst_bytes == 8) {
    APH_LAUNCH((k_cast<float, double>), grid_for(n), st, (const float*)src, (double*)dst, n);
  } else if (src_bytes == 8 && dst_bytes == 8) {
    return hipMemcpyAsync(dst, src, n * 8, hipMemcpyDeviceToDevice, st);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_assemble(const double* rho, const double* fk, const double* ffv,
                           const double* src, int nx, int ny, int nz, double cell, double dt,
                           int sign, double* expr, hipStream_t st) {
  const long long n = (long long)nx * ny * nz;
  if (n <= 0) return hipSuccess;
  APH_LAUNCH(k_assemble, grid_for(n), st, rho, fk, ffv, src, nx, ny, nz, cell, dt, sign, expr);
  return hipGetLastError();
}

hipError_t launch_flux_correct(const double* p, const double* rho, const double* fk,
                               const double* ffv, double* out, int nx, int ny, int nz,
                               double cell, double dt, hipStream_t st) {
  const long long n = (long long)nx * ny * nz;
  if (n <= 0) return hipSuccess;
  APH_LAUNCH(k_flux_correct, grid_for(n), st, p, rho, fk, ffv, out, nx, ny, nz, cell, dt);
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

namespace aph {

hipError_t launch_perturb(double* b, long long n, unsigned long long key, double amp,
                          double* partials, hipStream_t st) {
  int g = grid_for(n);
  if (g > 1024) g = 1024;
  APH_LAUNCH(k_perturb_stats, g, st, b, n, key, partials);
  APH_LAUNCH(k_fin_perturb_stats, 1, st, partials, g, n);
  APH_LAUNCH(k_perturb_apply, grid_for(n), st, b, n, key, amp, partials + 2 * g);
  return hipGetLastError();
}

}  // namespace aph
