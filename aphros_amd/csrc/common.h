// Shared device/host structures of the aphpois engine.
//
// Layout in HBM (per system, all SoA, x fastest, no halo slots; periodic
// neighbours are found by index wrap):
//   PCG   : kx, ky, kz  face conductances of the plus faces (type T)
//           S           1/sqrt(diag) scale (type T), m optional mass term (T)
//           r, p[2], q, xh  Krylov vectors of the scaled system (type T)
//           x64, b64    FP64 solution and right-hand side b = -e7
//           s4[4]       FP64 e0, e_xp, e_yp, e_zp (true residual)
//   CONJUGATE / JACOBI : f7[8] = e0..e7 (FP64), u, un/p[2], r, q (FP64)
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace aph {

constexpr int kBlock = 256;
constexpr int kXcd = 8;  // MI355X: 8 XCDs, workgroups dispatched round-robin

enum Method { kConjugate = 0, kJacobi = 1, kPcg = 2 };

// Per-system scalars, resident in device memory and owned by the kernels.
struct Scalars {
  double rho;         // (r,r) of the current iterate (scaled space for PCG)
  double rho_pend;    // (r,r) computed by the update kernel, pending a decision
  double alpha, beta;
  double rmax2;       // (r,r) at the last reliable update
  double gamma;       // stop metric / sqrt(rho) measured at the last update
  double tol, delta, h3;
  double bnorm2;      // (b,b) in FP64
  double relres, res_l2, res_max, residual;
  double mean_pending;  // mean to subtract from x64 (nullspace removal)
  double rz;            // (r, z) of the current iterate (multigrid mode)
  int iter, miniter, maxiter, done;
  int ru, converged, nru, status;
  int resmode, singular, first, cur;  // cur: index of the current p/u buffer
  int method, pc;  // pc: 1 if a multigrid preconditioner supplies z
  int xh_fold;     // PCG state of x^: 0 it holds the corrections since the
                   // last reliable update; 1 stale (that update folded x^ and
                   // alpha p into x64; the next CG operator leaves x^ alone);
                   // 2 stale, the next operator writes x^ = alpha p_in
  int zero_start;  // PCG: the start's x64 is zero-filled (the start's
                   // residual is b, with no accumulation pass); cleared by it
};

// One level of the multigrid preconditioner (level 0 = the fine grid).
// Operator of every level: (L u)_c = sum_faces k_f (u_c - u_nb) + m_c u_c.
struct MgLevel {
  int nx, ny, nz;
  int post;  // 1: both post-smoothing sweeps fused (k_mg_post); 2: pre- and
             // post-smoothing as one tile launch each (k_mg_tile); the
             // pre-smoothed iterate is then z[0], else (0) z[1]
  long long n;
  // coefficients: FP64 engines: plus-face conductances and D^-1 as separate
  // arrays (kx, ky, kz, dinv); FP32 engines: all four in bfloat16, packed by
  // groups of 4 cells (kc: kx[4], ky[4], kz[4], dinv[4] per group)
  const void* kx;
  const void* ky;
  const void* kz;
  const void* m;   // mass term (element type of the engine) or nullptr
  const void* dinv;
  const void* kc;
  void* r;         // right-hand side of the level (level 0: the CG residual)
  void* z[2];      // correction, ping-pong
  void* zs;        // first of two coarse corrections (levels below a W level)
};

constexpr int kMaxLevels = 12;

struct SysDev {
  int nx, ny, nz, has_m;
  long long n;
  // PCG operator and vectors (element type float or double, see precision)
  const void* kx;
  const void* ky;
  const void* kz;
  const void* S;
  const void* m;
  void* r;
  void* p[2];
  void* q;
  void* xh;
  double* x64;
  const double* b64;
  const double* s4[4];
  // reference-mode storage: component k of cell c at f7[k][c << f7sh]
  // (f7sh 0: eight SoA arrays; 3: f7[k] = expr + k of a dense AoS Expr)
  const double* f7[8];
  int f7sh;
  double* u[2];
  double* rr;
  double* pp[2];
  double* qq;
  // reductions
  double* partials;  // 4 * nblk_max doubles
  unsigned* counter;
  Scalars* sc;
  int neg;  // 1 if the diagonal is negative (A = -L): r^ = -S r64
  // 1 if every e0 equals -(e_xm + e_xp + e_ym + e_yp + e_zm + e_zp), summed
  // in that order, bit for bit (a zero row sum assembled face by face): the
  // FP64 residual then forms e0 from the neighbour coefficients it loads
  int e0_sum;
  // workgroups of this system in the stencil, streaming and residual
  // launches (= number of partials each finalize kernel reads)
  int nblk_st, nblk_sm, nblk_ru;
  // preconditioned mode (multigrid): the CG direction is built from z
  // (= M^-1 r) instead of r, and S is 1 (S == nullptr)
  int nlev;        // multigrid levels (0 = Jacobi-scaled CG without z)
  int mg_w;        // levels l < mg_w apply their coarse correction twice
  int mg_lb;       // first level of the bottom kernel (levels mg_lb..nlev-1)
  int nblk_pc;     // workgroups of the level-0 dot in the last smoothing sweep
  const void* z;   // vector the p update reads (r in Jacobi mode)
  const MgLevel* lev;  // device array of nlev levels
};

// One record per system in a launch: which blocks belong to it and how they
// tile the grid.  Streaming launches use tiles_x = nblk, tiles_y = tiles_z = 1.
struct Rec {
  int sys;
  int blk0;
  int tiles_x, tiles_y, tiles_z;
  int lx;  // threads per row (stencil), rows per block = kBlock / lx
  int bz;  // planes per block (stencil)
  int nblk;
  int ty;  // rows of threads of a whole-row tile (k_mg_post: output rows + 2 halo rows)
};

// Physical workgroup id -> logical id such that the workgroups that run on
// one XCD (ids x, x+8, x+16, ...) get a contiguous range of logical tiles and
// share halo rows in that XCD's L2.
__device__ __forceinline__ int xcd_remap(int b, int g) {
  const int per = g / kXcd, rem = g % kXcd;
  const int x = b % kXcd, local = b / kXcd;
  return x < rem ? x * (per + 1) + local : rem * (per + 1) + (x - rem) * per + local;
}

}  // namespace aph
