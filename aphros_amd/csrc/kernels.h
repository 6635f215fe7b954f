// Launch wrappers of the aphpois kernels (kernels.hip).
#pragma once

#include "common.h"

namespace aph {

// Damped Jacobi sweeps (weight 0.9, from zero) on the coarsest multigrid
// level.  Mixed solves to 1e-8, 40 / 20 / 12 sweeps: 512^3 and 256^3 the same
// iterations and time; with 32 spheres at 96^3, 160^3, 224^3, 320^3, 384^3:
// 48 / 45 / 45, 52 / 54 / 49, 36 / 34 / 36, 30 / 28 / 32, 27 / 27 / 27
// iterations, 20 the fastest or within 1 % at every size but 160^3 (12).
constexpr int kCoarseSweeps = 20;

// Device tables of one engine launch configuration.
struct LaunchSet {
  const SysDev* sds = nullptr;
  const Rec* rec_st = nullptr;  // stencil launch of the iteration kernel
  const int* b2r_st = nullptr;
  const Rec* rec_sm = nullptr;  // streaming launches
  const int* b2r_sm = nullptr;
  const Rec* rec_ru = nullptr;  // one-cell-per-thread stencil launches (FP64 residuals)
  const int* b2r_ru = nullptr;
  int nst = 0, nsm = 0, nru = 0;
  int vx_ru = 1;  // cells per thread of the PCG FP64 residual launch (rec_ru)
  int ns = 0;  // systems (one workgroup each in the finalize kernels)
  // multigrid preconditioner (pc = 1): per level l < nlev_max - 1, the
  // one-cell-per-thread tiling of level l (rec_lf) and of level l+1 (rec_lc)
  // over the systems that coarsen below l
  int pc = 0;
  int nlev_max = 0;
  int mg_w = 0;  // largest number of W levels over the systems
  int deep = 0;  // largest bottom level over the systems (no launches below)
  int bottom_at[kMaxLevels] = {};  // some system's bottom kernel starts here
  int start_only = 0;  // cycle launches for newly prepared systems only
  int coarse_sweeps = kCoarseSweeps;
  int coarse_lds = 0;  // dynamic LDS bytes of the bottom-levels kernel
  int vx_l[kMaxLevels] = {};  // cells per thread in x of the level-l sweeps
  const Rec* rec_lf[kMaxLevels] = {};
  const int* b2r_lf[kMaxLevels] = {};
  int nlf[kMaxLevels] = {};
  const Rec* rec_lc[kMaxLevels] = {};
  const int* b2r_lc[kMaxLevels] = {};
  int nlc[kMaxLevels] = {};
  // post-smoothing of level l: some systems fused (k_mg_post over the
  // whole-row tiles rec_lp, cells per thread vx_p), some with two sweeps
  // (the sweep launches skip the systems whose level has post = 1)
  int post_fused[kMaxLevels] = {};
  int post_plain[kMaxLevels] = {};
  int vx_p[kMaxLevels] = {};
  const Rec* rec_lp[kMaxLevels] = {};
  const int* b2r_lp[kMaxLevels] = {};
  int nlp[kMaxLevels] = {};
  // small levels whose pre- and post-smoothing run as one tile launch each
  // (k_mg_tile, MgLevel::post == 2; 8^3 cells per workgroup, tiles_x/y/z)
  int post_tile[kMaxLevels] = {};
  const Rec* rec_lt[kMaxLevels] = {};
  const int* b2r_lt[kMaxLevels] = {};
  int nlt[kMaxLevels] = {};
  // event-profiled cycles only: the launches below level 0 as one replay of
  // this graph (captured by capture_below0), nullptr: launched one by one
  hipGraphExec_t below0 = nullptr;
};

// threads of a k_mg_post workgroup
constexpr int kPostBlock = 1024;
// k_mg_tile: output cells per tile edge and threads per workgroup
constexpr int kTile = 8;
constexpr int kTileBlock = 256;

// LDS of the bottom-levels kernel, per level of n cells (element size es):
// z and a second buffer; with everything resident also the coefficients
// (FP32: bfloat16 packed by 4 cells; FP64: kx, ky, kz, D^-1), the mass term
// and the right-hand side.  Everything is resident when the sum over the
// bottom levels stays within kBottomLdsMax.
constexpr long long kBottomLdsMax = 156 * 1024;
__host__ __device__ inline long long lds_al16(long long b) { return (b + 15) & ~15LL; }
__host__ __device__ inline long long bottom_coef_bytes(long long n, int es) {
  return es == 4 ? (n + 3) / 4 * 32 : 4 * lds_al16(8 * n);
}
__host__ __device__ inline long long bottom_z_bytes(long long n, int es) {
  return 2 * lds_al16(n * es);
}
__host__ __device__ inline long long bottom_full_bytes(long long n, int es, int has_m) {
  return bottom_z_bytes(n, es) + lds_al16(bottom_coef_bytes(n, es)) +
         (has_m ? lds_al16(n * es) : 0) + lds_al16(n * es);
}

// kernels of one iteration, each reducing kernel followed by its finalize
// kernel: PCG: spmv, fin, update, fin, ru_accum, fin, ru_resid, fin and, with
// the multigrid preconditioner, [8] the V-cycle with its finalize;
// CONJUGATE: spmv, fin, update, fin; JACOBI: sweep, fin
int iter_kernels(int method, int pc = 0);
// kernels of one iteration in the captured graph: PCG without the multigrid
// preconditioner leaves the reliable-update kernels (4 .. 7) out; a system
// whose update triggers one then idles (its CG kernels return while ru is
// set) until the host launches them (launch_start) after the replay
int graph_kernels(int method, int pc = 0);
// Coefficients of the V-cycle levels: level 0 from the CG's conductances
// kx, ky, kz (converted to the levels' coefficient type: bfloat16 in FP32
// engines, where level 0 has arrays of its own; FP64 engines share them),
// coarse operators, inverse diagonals.
hipError_t launch_mg_setup(int fp32, const MgLevel* levels, int nlev, const void* kx,
                           const void* ky, const void* kz, hipStream_t st);
// The V-cycle with an event recorded before each launch (ev needs
// vcycle_launches(L) + 1 entries; kinds receives the launch kinds).
hipError_t launch_vcycle_profiled(int fp32, const LaunchSet& L, hipStream_t st, hipEvent_t* ev,
                                  int* kinds, int* nlaunch);
int vcycle_launches(const LaunchSet& L);
// Records into the capturing stream st the launches of a cycle below level 0
// (levels >= 1 with their W corrections), for LaunchSet::below0.
hipError_t capture_below0(int fp32, const LaunchSet& L, hipStream_t st);
hipError_t launch_kernel(int method, int fp32, int vx, const LaunchSet& L, hipStream_t st,
                         int j);
hipError_t launch_iter(int method, int fp32, int vx, const LaunchSet& L, hipStream_t st);
hipError_t launch_start(int method, int fp32, int vx, const LaunchSet& L, hipStream_t st);
hipError_t launch_true_res_f7(const LaunchSet& L, hipStream_t st);
hipError_t launch_check(const double* aos, long long start, long long ly, long long lz, int nx,
                        int ny, int nz, unsigned* flags, hipStream_t st);
hipError_t launch_convert_f7(const double* aos, long long start, long long ly, long long lz,
                             int nx, int ny, int nz, double* f7, int coefs_only, hipStream_t st);
hipError_t launch_convert_pcg(int fp32, const double* aos, long long start, long long ly,
                              long long lz, int nx, int ny, int nz, double sgn, void* kx,
                              void* ky, void* kz, void* S, void* m, double* s4, double* b,
                              int coefs_only, hipStream_t st);
hipError_t launch_convert_rhs(const double* src, long long start, long long ly, long long lz,
                              int nx, int ny, int nz, double scale, double* dst, hipStream_t st);
hipError_t launch_gather(const double* src, long long start, long long ly, long long lz, int nx,
                         int ny, int nz, double* dst, double sub, hipStream_t st);
hipError_t launch_scatter(const double* src, int nx, int ny, int nz, double* dst,
                          long long start, long long ly, long long lz, double sub,
                          hipStream_t st);
hipError_t launch_sumsq(const double* v, long long n, double* partials, unsigned* counter,
                        double* out, hipStream_t st);
hipError_t launch_fill(void* p, long long n, int elem_bytes, hipStream_t st);
// b2r[i] = 0 for i < n0, 1 for i < n0 + n1, else 2 (i < n0 + n1 + n2)
hipError_t launch_b2r3(int* b2r, int n0, int n1, int n2, hipStream_t st);
// x -= sub in place
hipError_t launch_shift(double* x, long long n, double sub, hipStream_t st);
// d = (x - sub) - xref: out[0] = mean of d (0 without remove_mean), out[1] =
// max |d - out[0]|, in one pass over x and xref; partials: 3 x 1024 doubles
hipError_t launch_max_diff(const double* x, const double* xref, long long n, double sub,
                           int remove_mean, double* partials, double* out, hipStream_t st);
// The projection's pressure system (dense FP64 Expr AoS, x fastest) from
// density (or face denominators fk, 3 n; one of rho and fk is null), face
// volume fluxes (3 n, lower faces) and source (may be null).
hipError_t launch_assemble(const double* rho, const double* fk, const double* ffv,
                           const double* src, int nx, int ny, int nz, double cell, double dt,
                           int sign, double* expr, hipStream_t st);
// Corrected face fluxes (3 n) for the solution p.
hipError_t launch_flux_correct(const double* p, const double* rho, const double* fk,
                               const double* ffv, double* out, int nx, int ny, int nz,
                               double cell, double dt, hipStream_t st);
// dst = src, n16 16-byte elements (plain streaming copy)
hipError_t launch_copy(const void* src, void* dst, long long n16, hipStream_t st);
// dst = (type of dst) src, element sizes 8 -> 4, 4 -> 8 or 8 -> 8 bytes
hipError_t launch_cast(const void* src, int src_bytes, void* dst, int dst_bytes, long long n,
                       hipStream_t st);
// b += amp * max|b| * (eta - mean eta), eta from the splitmix64 stream `key`;
// partials needs 2 * 1024 + 2 doubles
hipError_t launch_perturb(double* b, long long n, unsigned long long key, double amp,
                          double* partials, hipStream_t st);

}  // namespace aph
