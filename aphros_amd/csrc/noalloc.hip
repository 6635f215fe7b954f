// Allocation-free solve of a reference-mode system (CONJUGATE or JACOBI).
//
// The same kernels, tilings and reductions as aphpois_create +
// aphpois_set_system_device + aphpois_solve_device, with every buffer beyond
// the caller's Expr and solution in one file-scope device array (`scratch`):
// the system table, the scalars, the launch records, the block -> record map,
// the reduction partials, the check flags and the work vectors.  The Expr is
// read in place (component k of cell c at expr[8 c + k]).  No hipMalloc,
// hipFree, stream, event or graph is created.
//
// Each device holds its own copy of `scratch`; its address on a device comes
// from hipGetSymbolAddress, taken once per device with that device current.
// The name `scratch` itself is never handed to a kernel or a copy from host
// code (on the host it names a stand-in, not device memory).
//
// Calls that overlap on one device share that device's array and race: the
// caller serialises them (one stream, or one call at a time).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>
#include <string>

#include "engine.h"

namespace aph {

namespace {

// The largest system: 2^21 cells (128^3).  Work vectors: CONJUGATE r, p[2],
// q; JACOBI the second iterate (the first is the caller's x).
constexpr long long kNaMaxCells = 1LL << 21;
constexpr long long kNaMaxBlocks = 1LL << 16;  // records of one launch
constexpr long long kNaVecs = 4;

struct NaHeader {
  SysDev sd;
  Scalars sc;
  Rec rec[3];  // stencil, streaming, one-cell-per-thread launches
  unsigned flags;
};

constexpr long long al256(long long b) { return (b + 255) & ~255LL; }
constexpr long long kOfsHdr = 0;
constexpr long long kOfsB2r = al256(sizeof(NaHeader));
constexpr long long kOfsPart = kOfsB2r + al256(4 * kNaMaxBlocks);
constexpr long long kOfsVec = kOfsPart + al256(3 * 8 * kNaMaxBlocks);
constexpr long long kScratchBytes = kOfsVec + kNaVecs * 8 * kNaMaxCells;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

// Writes the header (by value) into the device array and clears the
// block -> record map (every launch has one record) and the partials.
__global__ __launch_bounds__(kBlock) void k_na_setup(NaHeader h, unsigned char* base,
                                                     int nb2r) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t == 0) *(NaHeader*)(base + kOfsHdr) = h;
  int* b2r = (int*)(base + kOfsB2r);
  for (int i = t; i < nb2r; i += gridDim.x * kBlock) b2r[i] = 0;
}

// x (n) <- second iterate when Jacobi ended on it
__global__ __launch_bounds__(kBlock) void k_na_copy(const double* __restrict__ src,
                                                    double* __restrict__ dst, long long n) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock)
    dst[i] = src[i];
}

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_ptr[kMaxDevices] = {};

// Address of `scratch` on the current device (taken once per device).
int scratch_ptr(int device, unsigned char** out) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_api_error(APHPOIS_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") +
                                                hipGetErrorString(e));
    g_ptr[device] = (unsigned char*)p;
  }
  *out = g_ptr[device];
  return APHPOIS_OK;
}

int grid_of(long long n) { return (int)std::max(1LL, std::min(8192LL, (n + kBlock - 1) / kBlock)); }

#define NA_HIP(expr)                                                                     \
  do {                                                                                   \
    const hipError_t e__ = (expr);                                                       \
    if (e__ != hipSuccess)                                                               \
      return set_api_error(APHPOIS_EDEVICE, std::string(#expr ": ") + hipGetErrorString(e__)); \
  } while (0)

int solve_noalloc(const aphpois_desc* d, const double* expr, const double* x0, double* x,
                  hipStream_t st, aphpois_info* info) {
  if (!d || !expr || !x) return set_api_error(APHPOIS_EINVAL, "NULL argument");
  if (d->method != APHPOIS_CONJUGATE && d->method != APHPOIS_JACOBI)
    return set_api_error(APHPOIS_EINVAL, "the no-allocation solve takes CONJUGATE or JACOBI");
  if (d->resmode != APHPOIS_RES_L2 && d->resmode != APHPOIS_RES_MAX)
    return set_api_error(APHPOIS_EINVAL, "resmode must be RES_L2 or RES_MAX");
  if (d->nx <= 0 || d->ny <= 0 || d->nz <= 0)
    return set_api_error(APHPOIS_EINVAL, "grid size must be positive");
  if (d->maxiter < 0) return set_api_error(APHPOIS_EINVAL, "maxiter < 0");
  const long long n = (long long)d->nx * d->ny * d->nz;
  if (n > kNaMaxCells)
    return set_api_error(APHPOIS_EINVAL,
                         "the no-allocation solve takes at most 2^21 cells (" +
                             std::to_string(n) + " given)");
  int ndev = 0;
  NA_HIP(hipGetDeviceCount(&ndev));
  if (d->device < 0 || d->device >= ndev || d->device >= kMaxDevices)
    return set_api_error(APHPOIS_EINVAL, "no such device");
  NA_HIP(hipSetDevice(d->device));
  unsigned char* base = nullptr;
  {
    const int rc = scratch_ptr(d->device, &base);
    if (rc != APHPOIS_OK) return rc;
  }
  const int method = d->method == APHPOIS_CONJUGATE ? kConjugate : kJacobi;
  Rec rst, rsm, rru;
  Tuning none;
  system_tiling(method, d->nx, d->ny, d->nz, 1, none, &rst, &rsm, &rru);
  const int nb = std::max(rst.nblk, std::max(rsm.nblk, rru.nblk));
  if (nb > kNaMaxBlocks)
    return set_api_error(APHPOIS_EINVAL, "launch tiling exceeds the no-allocation tables");
  rst.sys = rsm.sys = rru.sys = 0;
  rst.blk0 = rsm.blk0 = rru.blk0 = 0;

  NaHeader* dh = (NaHeader*)(base + kOfsHdr);
  double* vec = (double*)(base + kOfsVec);
  NaHeader h;
  std::memset(&h, 0, sizeof(h));
  SysDev& s = h.sd;
  s.nx = d->nx;
  s.ny = d->ny;
  s.nz = d->nz;
  s.n = n;
  for (int k = 0; k < 8; ++k) s.f7[k] = expr + k;
  s.f7sh = 3;
  s.u[0] = x;
  s.u[1] = vec;  // JACOBI
  s.rr = vec;    // CONJUGATE
  s.pp[0] = vec + n;
  s.pp[1] = vec + 2 * n;
  s.qq = vec + 3 * n;
  s.partials = (double*)(base + kOfsPart);
  s.sc = &dh->sc;
  s.nblk_st = rst.nblk;
  s.nblk_sm = rsm.nblk;
  s.nblk_ru = rru.nblk;
  Scalars& c = h.sc;
  c.tol = d->tol;
  c.h3 = d->h3;
  c.gamma = 1.0;
  c.miniter = d->miniter;
  c.maxiter = d->maxiter;
  c.resmode = d->resmode;
  c.first = 1;
  c.method = method;
  h.rec[0] = rst;
  h.rec[1] = rsm;
  h.rec[2] = rru;

  const auto t0 = std::chrono::steady_clock::now();
  hipLaunchKernelGGL(k_na_setup, dim3(grid_of(nb)), dim3(kBlock), 0, st, h, base, nb);
  NA_HIP(hipGetLastError());
  NA_HIP(launch_check(expr, 0, d->nx, (long long)d->nx * d->ny, d->nx, d->ny, d->nz,
                      &dh->flags, st));
  unsigned flags = 0;
  NA_HIP(hipMemcpyAsync(&flags, &dh->flags, sizeof(flags), hipMemcpyDeviceToHost, st));
  NA_HIP(hipStreamSynchronize(st));
  if (flags & 1u) return set_api_error(APHPOIS_ENUMERIC, "system contains NaN or Inf");
  if (method == kJacobi && (flags & 4u))
    return set_api_error(APHPOIS_EINVAL, "zero diagonal coefficient (Jacobi divides by e0)");
  // initial guess in x
  if (!x0) {
    NA_HIP(launch_fill(x, n, 8, st));
  } else if (x0 != x) {
    hipLaunchKernelGGL(k_na_copy, dim3(grid_of(n)), dim3(kBlock), 0, st, x0, x, n);
    NA_HIP(hipGetLastError());
  }

  LaunchSet L;
  L.sds = &dh->sd;
  const int* b2r = (const int*)(base + kOfsB2r);
  L.rec_st = &dh->rec[0];
  L.rec_sm = &dh->rec[1];
  L.rec_ru = &dh->rec[2];
  L.b2r_st = L.b2r_sm = L.b2r_ru = b2r;
  L.nst = rst.nblk;
  L.nsm = rsm.nblk;
  L.nru = rru.nblk;
  L.ns = 1;
  NA_HIP(launch_start(method, 0, 1, L, st));
  // iterations between polls of the done flag; launches after the stop are
  // no-ops (every kernel returns on done), so the result does not depend on it
  const int k = 16;
  const int maxloops = (std::max(d->maxiter, d->miniter) + 1) / k + 4;
  Scalars hs;
  for (int loop = 0;; ++loop) {
    for (int i = 0; i < k; ++i) {
      for (int j = 0; j < iter_kernels(method); ++j) NA_HIP(launch_kernel(method, 0, 1, L, st, j));
    }
    NA_HIP(hipMemcpyAsync(&hs, &dh->sc, sizeof(hs), hipMemcpyDeviceToHost, st));
    NA_HIP(hipStreamSynchronize(st));
    if (hs.done) break;
    if (loop > maxloops)
      return set_api_error(APHPOIS_EDEVICE, "solver did not terminate within maxiter");
  }
  NA_HIP(launch_true_res_f7(L, st));
  if (method == kJacobi && hs.cur == 1) {
    hipLaunchKernelGGL(k_na_copy, dim3(grid_of(n)), dim3(kBlock), 0, st, vec, x, n);
    NA_HIP(hipGetLastError());
  }
  NA_HIP(hipMemcpyAsync(&hs, &dh->sc, sizeof(hs), hipMemcpyDeviceToHost, st));
  NA_HIP(hipStreamSynchronize(st));
  const double sec =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (info) {
    std::memset(info, 0, sizeof(*info));
    info->iter = hs.iter;
    info->residual = hs.residual;
    info->res_l2 = hs.res_l2;
    info->res_max = hs.res_max;
    info->true_relres = hs.relres;
    info->converged = hs.converged;
    info->status = hs.status;
    info->seconds = sec;
    info->max_diff = -1.0;
  }
  return APHPOIS_OK;
}

}  // namespace

}  // namespace aph

extern "C" {

int aphpois_solve_device_noalloc(const aphpois_desc* desc, const double* expr_dev,
                                 const double* x0_dev, double* x_dev, void* stream,
                                 aphpois_info* info) {
  return aph::solve_noalloc(desc, expr_dev, x0_dev, x_dev, (hipStream_t)stream, info);
}

long long aphpois_noalloc_max_cells(void) { return aph::kNaMaxCells; }

}  // extern "C"
