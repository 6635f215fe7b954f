// Host engine of aphpois (see engine.h).
#include "engine.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace aph {

namespace {

int pow2floor(int v) {
  int p = 1;
  while (p * 2 <= v) p *= 2;
  return p;
}

long long cdiv(long long a, long long b) { return (a + b - 1) / b; }

// Stencil tiling of one system: lx lanes of VX cells per row (at most
// lx_max by default), kBlock/lx rows, bz planes per workgroup.
Rec make_stencil_rec(int nx, int ny, int nz, int vx, int blocks_target, const Tuning& tu,
                     int lx_max = 64) {
  Rec r{};
  const int cols = std::max(1, nx / vx);
  int lx = tu.lx > 0 ? tu.lx : std::min(lx_max, pow2floor(cols));
  lx = std::max(1, std::min(lx, kBlock));
  const int rows = kBlock / lx;
  r.lx = lx;
  r.tiles_x = (int)cdiv(cols, lx);
  r.tiles_y = (int)cdiv(ny, rows);
  int bz;
  if (tu.bz > 0) {
    bz = tu.bz;
  } else {
    const long long cols_tiles = (long long)r.tiles_x * r.tiles_y;
    const long long tz = std::max(1LL, cdiv(blocks_target, cols_tiles));
    bz = (int)cdiv(nz, tz);
  }
  bz = std::max(1, std::min(bz, nz));
  r.bz = bz;
  r.tiles_z = (int)cdiv(nz, bz);
  r.nblk = r.tiles_x * r.tiles_y * r.tiles_z;
  return r;
}

// Fused pre-smoothing + restriction (k_mg_pre) and post-smoothing (k_mg_post)
// of a level: whole-row tiles of lx = nx / vx threads per row and
// ty = min(kPostBlock / lx, ny + 2) rows of threads (even; ty - 2 output
// rows), bz planes (even).  Only for the large levels (n >= 2^24 cells: the
// march is long enough to hide the halo rows and planes it recomputes; such a
// level coarsens, so its edges are even) whose rows fit 1/8 of the workgroup
// (ty >= 8 rows).
constexpr long long kFuseMinCells = 1LL << 24;
// Levels below level 0 of FP32 engines with at most this many cells and
// nx % 4 == 0 smooth in tile launches (k_mg_tile: one launch and one round of
// global loads per pre- or post-smoothing instead of two z-marching
// launches).  Per visit at 64^3: 8.1 + 8.6 us against 12 + 11 us for the
// sweep pairs; at 32^3 6 + 6 us against 11.5 + 11.5 us; at 128^3 the tiles'
// halos (4.5x the level's bytes) miss L2 and the tiles took 37 + 38 us
// against 22 + 22 us, so 128^3 keeps the sweeps.
constexpr long long kTileMaxCells = 1LL << 18;

bool make_post_rec(int nx, int ny, int nz, int vx, int ncu, Rec* out) {
  const long long n = (long long)nx * ny * nz;
  if (vx <= 0 || nx % vx != 0 || n < kFuseMinCells) return false;
  const int lx = nx / vx;
  if (lx > kPostBlock / 8) return false;
  Rec r{};
  r.lx = lx;
  // even: the fused restriction pairs output rows from an even row
  r.ty = std::min(kPostBlock / lx, ny + 2) & ~1;
  r.tiles_x = 1;
  r.tiles_y = (int)cdiv(ny, r.ty - 2);
  // One workgroup is resident per CU (1024 threads at 128 VGPRs per lane) and
  // a workgroup's time grows with the planes it marches, bz plus two halo
  // planes.  The z split minimises rounds * (bz + 2), rounds = workgroups / CUs
  // rounded up; of equal costs the one with fewer workgroups (less halo) wins.
  // bz is even: the restriction pairs planes (nz is even on these levels).
  const long long cus = std::max(1, ncu);
  long long best = -1;
  for (int tz = 1; tz <= nz; ++tz) {
    int bz = (int)cdiv(nz, tz);
    bz = std::min(bz + (bz & 1), nz);
    const int tiles_z = (int)cdiv(nz, bz);
    if (tiles_z != tz) continue;
    const long long cost = cdiv((long long)r.tiles_y * tiles_z, cus) * (bz + 2);
    if (best < 0 || cost < best) {
      best = cost;
      r.bz = bz;
      r.tiles_z = tiles_z;
    }
  }
  r.nblk = r.tiles_y * r.tiles_z;
  *out = r;
  return true;
}

}  // namespace

void system_tiling(int method, int nx, int ny, int nz, int vx, const Tuning& tu, Rec* st,
                   Rec* sm, Rec* ru, int vx_ru, int fp64) {
  // The tiling of a system depends on that system only, so its reduction
  // trees (and results) are the same alone and inside a batch.
  // ~one workgroup per 2048 cells, at most 1024 (about one resident wave
  // set of the 256 CUs: long z marches, few re-read halo planes)
  const long long n = (long long)nx * ny * nz;
  const long long tgt = std::min(1024LL, std::max(1LL, cdiv(n, 2048)));
  // the FP64 PCG operator: rows of up to 128 lanes (whole rows at 256^3:
  // 0.249 -> 0.240 ms per launch against 64 lanes).  The FP32 operator keeps
  // 64: whole rows at 512^3 took 0.718 against 0.722 ms but fetched 3.51
  // instead of 3.15 GB per launch (two output rows per block re-read two halo
  // rows of z and p)
  *st = make_stencil_rec(nx, ny, nz, method == kPcg ? vx : 1, tgt, tu,
                         method == kPcg && fp64 ? 128 : 64);
  Tuning none;
  *ru = make_stencil_rec(nx, ny, nz, vx_ru, tgt, none);
  Rec b{};
  const int cells = tu.sm > 0 ? tu.sm : (method == kPcg ? 4 * vx : 4);
  b.nblk = (int)std::min(16384LL, std::max(1LL, cdiv(n, (long long)kBlock * cells)));
  b.tiles_x = b.nblk;
  b.tiles_y = b.tiles_z = 1;
  b.lx = kBlock;
  b.bz = 1;
  *sm = b;
}

int Engine::fail(int code, const std::string& msg) {
  err_ = msg;
  return code;
}

int Engine::hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return APHPOIS_OK;
  err_ = std::string(what) + ": " + hipGetErrorString(e);
  return e == hipErrorOutOfMemory ? APHPOIS_ENOMEM : APHPOIS_EDEVICE;
}

#define APH_TRY(expr)                 \
  do {                                \
    const int rc__ = (expr);          \
    if (rc__ != APHPOIS_OK) return rc__; \
  } while (0)
#define APH_HIP(expr) APH_TRY(hip((expr), #expr))

Engine::~Engine() {
  // teardown: errors cannot be reported any more and are ignored
  if (st_) (void)hipStreamSynchronize(st_);
  for (auto& s : slots_) {
    if (s) free_system(*s);
  }
  if (gexec_) (void)hipGraphExecDestroy(gexec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  if (below_exec_) (void)hipGraphExecDestroy(below_exec_);
  if (below_graph_) (void)hipGraphDestroy(below_graph_);
  if (d_sc_) (void)hipFree(d_sc_);
  if (h_sc_) (void)hipHostFree(h_sc_);
  if (d_counter_) (void)hipFree(d_counter_);
  if (d_sds_) (void)hipFree(d_sds_);
  if (d_rec_) (void)hipFree(d_rec_);
  if (d_b2r_) (void)hipFree(d_b2r_);
  if (staging_) (void)hipFree(staging_);
  release(staging_bytes_);
  if (d_flags_) (void)hipFree(d_flags_);
  if (d_red_) (void)hipFree(d_red_);
  if (d_red_counter_) (void)hipFree(d_red_counter_);
  if (ev0_) (void)hipEventDestroy(ev0_);
  if (ev1_) (void)hipEventDestroy(ev1_);
  for (hipEvent_t e : pev_) (void)hipEventDestroy(e);
  for (hipEvent_t e : pev_mg_) (void)hipEventDestroy(e);
  if (st_) (void)hipStreamDestroy(st_);
}

int Engine::init(int device, int method, int precision, int resmode, double tol, int miniter,
                 int maxiter, int nullspace, double delta, int check_every, long long budget,
                 int capacity, int precond, int mg_w) {
  if (method < 0 || method > 2) return fail(APHPOIS_EINVAL, "unknown method");
  if (precision < 0 || precision > 1) return fail(APHPOIS_EINVAL, "unknown precision");
  if (resmode < 0 || resmode > 2) return fail(APHPOIS_EINVAL, "unknown resmode");
  if (resmode == 2 && method != kPcg)
    return fail(APHPOIS_EINVAL, "RES_REL_TRUE needs method PCG");
  if (precision == 1 && method != kPcg)
    return fail(APHPOIS_EINVAL, "MIXED precision needs method PCG");
  if (maxiter < 0) return fail(APHPOIS_EINVAL, "maxiter < 0");
  if (capacity <= 0) capacity = 1024;
  int ndev = 0;
  APH_HIP(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(APHPOIS_EINVAL, "no such device");
  device_ = device;
  APH_HIP(hipSetDevice(device_));
  APH_HIP(hipDeviceGetAttribute(&ncu_, hipDeviceAttributeMultiprocessorCount, device_));
  method_ = method;
  fp32_ = precision == 1;
  if (precond < 0 || precond > 1) return fail(APHPOIS_EINVAL, "unknown preconditioner");
  if (precond == 1 && method != kPcg)
    return fail(APHPOIS_EINVAL, "the multigrid preconditioner needs method PCG");
  pc_ = precond == 1;
  if (mg_w < APHPOIS_MG_W_AUTO || mg_w > kMaxLevels)
    return fail(APHPOIS_EINVAL, "mg_w out of range");
  mg_w_ = mg_w;
  resmode_ = resmode;
  tol_ = tol;
  miniter_ = miniter;
  maxiter_ = maxiter;
  nullspace_ = nullspace;
  // fewer reliable updates for the same iteration count: 512^3 mixed, 35
  // iterations at 0.1 and 0.05 (0.166 / 0.160 s), 36 at 0.03 and 0.02
  delta_ = delta > 0 ? delta : 0.05;
  check_every_ = check_every;
  budget_ = budget;
  capacity_ = capacity;
  APH_HIP(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
  APH_HIP(hipMalloc(&d_sc_, sizeof(Scalars) * capacity_));
  APH_HIP(hipHostMalloc(&h_sc_, sizeof(Scalars) * capacity_, hipHostMallocDefault));
  std::memset(h_sc_, 0, sizeof(Scalars) * capacity_);
  APH_HIP(hipMalloc(&d_counter_, sizeof(unsigned) * capacity_));
  APH_HIP(hipMemset(d_counter_, 0, sizeof(unsigned) * capacity_));
  APH_HIP(hipMemset(d_sc_, 0, sizeof(Scalars) * capacity_));
  APH_HIP(hipMalloc(&d_flags_, sizeof(unsigned)));
  // reduction scratch of k_sumsq (1024 partials + result), of the RHS
  // perturbation (2 x 1024 partials + 2 results) and of max_diff (3 x 1024
  // partials + 2 results)
  APH_HIP(hipMalloc(&d_red_, sizeof(double) * 3080));
  APH_HIP(hipMalloc(&d_red_counter_, sizeof(unsigned)));
  APH_HIP(hipMemset(d_red_counter_, 0, sizeof(unsigned)));
  APH_HIP(hipEventCreate(&ev0_));
  APH_HIP(hipEventCreate(&ev1_));
  return APHPOIS_OK;
}

long long Engine::bytes() const {
  long long b = staging_bytes_;
  for (auto& s : slots_) {
    if (s) b += s->bytes;
  }
  return b;
}

bool Engine::reserve(long long nbytes) {
  if (shared_) {
    long long cur = shared_->load();
    do {
      if (budget_ > 0 && cur + nbytes > budget_) return false;
    } while (!shared_->compare_exchange_weak(cur, cur + nbytes));
    return true;
  }
  return budget_ <= 0 || bytes() + nbytes <= budget_;
}

void Engine::release(long long nbytes) {
  if (shared_) shared_->fetch_sub(nbytes);
}

int Engine::alloc(System& s, void** p, long long nbytes) {
  if (!reserve(nbytes)) {
    return fail(APHPOIS_EBUDGET, "memory budget exceeded (" +
                                     std::to_string((shared_ ? shared_->load() : bytes()) +
                                                    nbytes) +
                                     " > " + std::to_string(budget_) + " bytes)");
  }
  hipError_t e = hipMalloc(p, (size_t)std::max(nbytes, 256LL));
  if (e != hipSuccess) {
    *p = nullptr;
    release(nbytes);
    return hip(e, "hipMalloc");
  }
  s.allocs.push_back(*p);
  s.bytes += nbytes;
  return APHPOIS_OK;
}

void Engine::free_system(System& s) {
  for (void* p : s.allocs) (void)hipFree(p);
  s.allocs.clear();
  release(s.bytes);
  s.bytes = 0;
  if (s.partials) (void)hipFree(s.partials);
  s.partials = nullptr;
  s.partials_cap = 0;
}

int Engine::add_system(int nx, int ny, int nz, double h3, int* slot) {
  if (nx < 1 || ny < 1 || nz < 1) return fail(APHPOIS_EINVAL, "grid size must be >= 1");
  if ((long long)nx * ny * nz > (1LL << 40)) return fail(APHPOIS_EINVAL, "grid too large");
  if (!(h3 > 0)) return fail(APHPOIS_EINVAL, "h3 must be > 0");
  int id = -1;
  for (int i = 0; i < (int)slots_.size(); ++i) {
    if (!slots_[i]) {
      id = i;
      break;
    }
  }
  if (id < 0) {
    if ((int)slots_.size() >= capacity_) return fail(APHPOIS_EBUDGET, "no free system slot");
    id = (int)slots_.size();
    slots_.emplace_back();
  }
  auto s = std::make_unique<System>();
  s->nx = nx;
  s->ny = ny;
  s->nz = nz;
  s->n = (long long)nx * ny * nz;
  s->h3 = h3;
  const long long n = s->n;
  // storage (everything except the optional mass term, added at set_system)
  const int es = fp32_ ? 4 : 8;
  System& S = *s;
  S.dev.nx = nx;
  S.dev.ny = ny;
  S.dev.nz = nz;
  S.dev.n = n;
  S.dev.sc = d_sc_ + id;
  S.dev.counter = d_counter_ + id;
  void* p = nullptr;
  int rc = APHPOIS_OK;
  auto A = [&](void** dst, long long b) {
    if (rc == APHPOIS_OK) rc = alloc(S, dst, b);
  };
  if (method_ == kPcg) {
    void *kx, *ky, *kz, *Sc = nullptr, *r, *p0, *p1, *q, *xh, *x64, *b64, *s4;
    if (pc_) {
      // multigrid levels: halve while every edge is even and the coarse grid
      // keeps at least 8 cells; the coarsest level must fit the LDS solver
      S.lev.clear();
      int lx = nx, ly = ny, lz = nz;
      for (;;) {
        MgLevel L{};
        L.nx = lx;
        L.ny = ly;
        L.nz = lz;
        L.n = (long long)lx * ly * lz;
        S.lev.push_back(L);
        if ((int)S.lev.size() >= kMaxLevels || (lx | ly | lz) & 1 ||
            (long long)(lx / 2) * (ly / 2) * (lz / 2) < 8)
          break;
        lx /= 2;
        ly /= 2;
        lz /= 2;
      }
      const long long cmax = fp32_ ? 8192 : 4096;
      if (S.lev.size() < 2 || S.lev.back().n > cmax) {
        slots_[id].reset();
        return fail(APHPOIS_EINVAL,
                    "grid cannot be coarsened to a small enough coarsest level (edges need "
                    "enough factors of 2) for the multigrid preconditioner");
      }
    }
    A(&kx, n * es);
    A(&ky, n * es);
    A(&kz, n * es);
    if (!pc_) A(&Sc, n * es);
    A(&r, n * es);
    A(&p0, n * es);
    A(&p1, n * es);
    A(&q, n * es);
    A(&xh, n * es);
    A(&x64, n * 8);
    A(&b64, n * 8);
    A(&s4, n * 32);
    if (rc != APHPOIS_OK) {
      free_system(S);
      slots_[id].reset();
      return rc;
    }
    S.dev.kx = kx;
    S.dev.ky = ky;
    S.dev.kz = kz;
    S.dev.S = Sc;
    S.dev.r = r;
    S.dev.p[0] = p0;
    S.dev.p[1] = p1;
    S.dev.q = q;
    S.dev.xh = xh;
    S.dev.x64 = (double*)x64;
    S.dev.b64 = (const double*)b64;
    for (int k = 0; k < 4; ++k) S.dev.s4[k] = (const double*)s4 + k * n;
    S.dev.z = r;
    S.dev.nlev = 0;
    if (pc_) {
      // level 0 shares the residual of the CG and, in FP64, its conductances;
      // the levels of an FP32 engine store kx, ky, kz, D^-1 in bfloat16,
      // packed by groups of 4 cells (8 bytes per cell)
      for (size_t l = 0; l < S.lev.size() && rc == APHPOIS_OK; ++l) {
        MgLevel& L = S.lev[l];
        void *z0, *z1;
        if (fp32_) {
          void* kc = nullptr;
          A(&kc, (L.n + 3) / 4 * 32);
          L.kc = kc;
        } else {
          void* dinv = nullptr;
          if (l == 0) {
            L.kx = kx;
            L.ky = ky;
            L.kz = kz;
          } else {
            void *ckx, *cky, *ckz;
            A(&ckx, L.n * es);
            A(&cky, L.n * es);
            A(&ckz, L.n * es);
            L.kx = ckx;
            L.ky = cky;
            L.kz = ckz;
          }
          A(&dinv, L.n * es);
          L.dinv = dinv;
        }
        if (l == 0) {
          L.r = r;
        } else {
          void* cr;
          A(&cr, L.n * es);
          L.r = cr;
        }
        A(&z0, L.n * es);
        A(&z1, L.n * es);
        L.z[0] = z0;
        L.z[1] = z1;
      }
      // W levels l < w (their coarse level l + 1 below the coarsest) keep the
      // first of two coarse corrections in zs of level l + 1.  Automatic:
      // by cell count (mixed solves to 1e-8 of three sphere geometries at
      // 128^3 .. 512^3: w = 0 fastest up to ~160^3, 1 or 2 within a few
      // percent from 192^3 to 320^3, 3 from 384^3)
      int wsel = mg_w_;
      if (wsel < 0) {
        const double lg = std::log2((double)S.n);
        wsel = lg < 22.5 ? 0 : (lg < 24.5 ? 1 : (lg < 25.5 ? 2 : 3));
      }
      const int w = std::max(0, std::min(wsel, (int)S.lev.size() - 2));
      S.dev.mg_w = w;
      // the bottom kernel takes the levels from the first one below every W
      // level with at most 4096 (FP32) / 2048 (FP64) cells, or the coarsest
      // alone: at most 64 KiB of LDS for z and a second buffer per level
      const long long bcap = fp32_ ? 4096 : 2048;
      int lb = (int)S.lev.size() - 1;
      for (int l = std::max(1, w + 1); l + 1 < (int)S.lev.size(); ++l) {
        if (S.lev[l].n <= bcap) {
          lb = l;
          break;
        }
      }
      S.dev.mg_lb = lb;
      // levels above the bottom kernel that fuse their post-smoothing (a
      // property of the system alone, like all of its tiling)
      for (int l = 0; l < lb; ++l) {
        Rec t;
        MgLevel& L = S.lev[l];
        L.post = make_post_rec(L.nx, L.ny, L.nz, fp32_ ? 4 : 2, ncu_, &t)
                     ? 1
                     : (fp32_ && l >= 1 && L.n <= kTileMaxCells && L.nx % 4 == 0 ? 2 : 0);
      }
      for (int l = 1; l <= w && rc == APHPOIS_OK; ++l) {
        // a fused or tile level's second cycle leaves z[1] (the first
        // correction) alone until its last sweep, which adds it in place;
        // plain levels pre-smooth into z[1] and keep a copy in zs
        if (S.lev[l].post != 0) {
          S.lev[l].zs = S.lev[l].z[1];
          continue;
        }
        void* zs = nullptr;
        A(&zs, S.lev[l].n * es);
        S.lev[l].zs = zs;
      }
      void* dl = nullptr;
      A(&dl, (long long)sizeof(MgLevel) * (long long)S.lev.size());
      if (rc != APHPOIS_OK) {
        free_system(S);
        slots_[id].reset();
        return rc;
      }
      S.d_lev = (MgLevel*)dl;
      S.dev.nlev = (int)S.lev.size();
      S.dev.lev = S.d_lev;
      S.dev.z = S.lev[0].z[1];  // level-0 correction after the post-smoothing
    }
  } else {
    void* f7;
    A(&f7, n * 64);
    void *u0 = nullptr, *u1 = nullptr, *rr = nullptr, *p0 = nullptr, *p1 = nullptr, *q = nullptr;
    A(&u0, n * 8);
    if (method_ == kJacobi) {
      A(&u1, n * 8);
    } else {
      A(&rr, n * 8);
      A(&p0, n * 8);
      A(&p1, n * 8);
      A(&q, n * 8);
    }
    if (rc != APHPOIS_OK) {
      free_system(S);
      slots_[id].reset();
      return rc;
    }
    for (int k = 0; k < 8; ++k) S.dev.f7[k] = (const double*)f7 + k * n;
    S.dev.u[0] = (double*)u0;
    S.dev.u[1] = method_ == kJacobi ? (double*)u1 : (double*)u0;
    S.dev.rr = (double*)rr;
    S.dev.pp[0] = (double*)p0;
    S.dev.pp[1] = (double*)p1;
    S.dev.qq = (double*)q;
  }
  (void)p;
  slots_[id] = std::move(s);
  dirty_ = true;
  *slot = id;
  return APHPOIS_OK;
}

int Engine::remove(int slot) {
  System* s = sys(slot);
  if (!s) return fail(APHPOIS_EINVAL, "no such slot");
  APH_HIP(hipStreamSynchronize(st_));
  free_system(*s);
  slots_[slot].reset();
  dirty_ = true;
  return APHPOIS_OK;
}

// Copies the strided range of a host or device field into the staging buffer;
// *dptr then addresses slot 0 of the range (use start' = 0).
int Engine::upload_strided(const double* src, bool on_device, long long start, long long ly,
                           long long lz, int nx, int ny, int nz, int comps,
                           const double** dptr) {
  if (start < 0 || ly < nx || lz < ly * (long long)ny)
    return fail(APHPOIS_EINVAL, "bad strides (need lead_y >= nx, lead_z >= lead_y*ny)");
  const long long last = start + (long long)(nz - 1) * lz + (long long)(ny - 1) * ly + (nx - 1);
  const long long count = (last - start + 1) * comps;
  const long long need = count * 8;
  APH_TRY(ensure_staging(need));
  APH_HIP(hipMemcpyAsync(staging_, src + start * comps, need,
                         on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st_));
  *dptr = (const double*)staging_;
  return APHPOIS_OK;
}

int Engine::set_system(int slot, const double* expr, bool on_device, long long start,
                       long long ly, long long lz) {
  System* sp = sys(slot);
  if (!sp) return fail(APHPOIS_EINVAL, "no such slot");
  if (!expr) return fail(APHPOIS_EINVAL, "expr is NULL");
  System& s = *sp;
  const double* d = nullptr;
  APH_TRY(upload_strided(expr, on_device, start, ly, lz, s.nx, s.ny, s.nz, 8, &d));
  return set_system_staged(slot, d, ly, lz);
}

int Engine::ensure_staging(long long need) {
  if (need <= staging_bytes_) return APHPOIS_OK;
  if (staging_) {
    APH_HIP(hipStreamSynchronize(st_));
    APH_HIP(hipFree(staging_));
    staging_ = nullptr;
    release(staging_bytes_);
    staging_bytes_ = 0;
  }
  if (!reserve(need)) return fail(APHPOIS_EBUDGET, "memory budget exceeded by the staging buffer");
  const hipError_t e = hipMalloc(&staging_, need);
  if (e != hipSuccess) {
    staging_ = nullptr;
    release(need);
    return hip(e, "hipMalloc (staging)");
  }
  staging_bytes_ = need;
  return APHPOIS_OK;
}

int Engine::assemble(int slot, const double* rho, const double* fk, const double* ffv,
                     const double* src, double cell, double dt, int sign, bool on_device) {
  System* sp = sys(slot);
  if (!sp) return fail(APHPOIS_EINVAL, "no such slot");
  if ((!rho) == (!fk) || !ffv) return fail(APHPOIS_EINVAL, "need ffv and one of rho, fk");
  if (!(cell > 0) || !(dt > 0) || (sign != 1 && sign != -1))
    return fail(APHPOIS_EINVAL, "need cell > 0, dt > 0 and sign +1 or -1");
  System& s = *sp;
  const long long n = s.n;
  APH_TRY(ensure_staging(n * 8 * 8));
  const double *drho = rho, *dfk = fk, *dffv = ffv, *dsrc = src;
  void* tmp = nullptr;
  long long tmp_bytes = 0;
  if (!on_device) {
    // inputs through a temporary device buffer: density (n) or face
    // denominators (3 n), face fluxes (3 n), source (n)
    const long long nden = rho ? n : 3 * n;
    const long long need = 8 * (nden + 3 * n + (src ? n : 0));
    if (!reserve(need))
      return fail(APHPOIS_EBUDGET, "memory budget exceeded by the assembly inputs");
    const hipError_t ea = hipMalloc(&tmp, need);
    if (ea != hipSuccess) {
      release(need);
      return hip(ea, "hipMalloc (assembly inputs)");
    }
    tmp_bytes = need;
    double* t = (double*)tmp;
    hipError_t e =
        hipMemcpyAsync(t, rho ? rho : fk, nden * 8, hipMemcpyHostToDevice, st_);
    if (e == hipSuccess)
      e = hipMemcpyAsync(t + nden, ffv, 3 * n * 8, hipMemcpyHostToDevice, st_);
    if (e == hipSuccess && src)
      e = hipMemcpyAsync(t + nden + 3 * n, src, n * 8, hipMemcpyHostToDevice, st_);
    if (e != hipSuccess) {
      (void)hipStreamSynchronize(st_);
      (void)hipFree(tmp);
      release(need);
      return hip(e, "assemble upload");
    }
    drho = rho ? t : nullptr;
    dfk = fk ? t : nullptr;
    dffv = t + nden;
    dsrc = src ? t + nden + 3 * n : nullptr;
  }
  hipError_t e = launch_assemble(drho, dfk, dffv, dsrc, s.nx, s.ny, s.nz, cell, dt, sign,
                                 (double*)staging_, st_);
  if (tmp) {
    const hipError_t e2 = hipStreamSynchronize(st_);
    (void)hipFree(tmp);
    release(tmp_bytes);
    if (e == hipSuccess) e = e2;
  }
  APH_HIP(e);
  return set_system_staged(slot, (const double*)staging_, s.nx, (long long)s.nx * s.ny);
}

int Engine::flux_correct(int slot, const double* p, const double* rho, const double* fk,
                         const double* ffv, double* out, double cell, double dt,
                         bool on_device) {
  System* sp = sys(slot);
  if (!sp) return fail(APHPOIS_EINVAL, "no such slot");
  if (!p || (!rho) == (!fk) || !ffv || !out)
    return fail(APHPOIS_EINVAL, "need p, ffv, out and one of rho, fk");
  if (!(cell > 0) || !(dt > 0)) return fail(APHPOIS_EINVAL, "need cell > 0 and dt > 0");
  const long long n = sp->n;
  if (on_device) {
    APH_HIP(launch_flux_correct(p, rho, fk, ffv, out, sp->nx, sp->ny, sp->nz, cell, dt, st_));
    APH_HIP(hipStreamSynchronize(st_));
    return APHPOIS_OK;
  }
  // p (n), density (n) or face denominators (3 n), fluxes (3 n) in, 3 n out
  const long long nden = rho ? n : 3 * n;
  APH_TRY(ensure_staging(8 * (n + nden + 6 * n)));
  double* t = (double*)staging_;
  double* tden = t + n;
  double* tffv = tden + nden;
  double* tout = tffv + 3 * n;
  APH_HIP(hipMemcpyAsync(t, p, n * 8, hipMemcpyHostToDevice, st_));
  APH_HIP(hipMemcpyAsync(tden, rho ? rho : fk, nden * 8, hipMemcpyHostToDevice, st_));
  APH_HIP(hipMemcpyAsync(tffv, ffv, 3 * n * 8, hipMemcpyHostToDevice, st_));
  APH_HIP(launch_flux_correct(t, rho ? tden : nullptr, fk ? tden : nullptr, tffv, tout, sp->nx,
                              sp->ny, sp->nz, cell, dt, st_));
  APH_HIP(hipMemcpyAsync(out, tout, 3 * n * 8, hipMemcpyDeviceToHost, st_));
  APH_HIP(hipStreamSynchronize(st_));
  return APHPOIS_OK;
}

int Engine::set_system_staged(int slot, const double* d, long long ly, long long lz) {
  System& s = *sys(slot);
  s.bnorm2 = -1;
  APH_HIP(hipMemsetAsync(d_flags_, 0, sizeof(unsigned), st_));
  APH_HIP(launch_check(d, 0, ly, lz, s.nx, s.ny, s.nz, d_flags_, st_));
  unsigned flags = 0;
  APH_HIP(hipMemcpyAsync(&flags, d_flags_, sizeof(unsigned), hipMemcpyDeviceToHost, st_));
  APH_HIP(hipStreamSynchronize(st_));
  if (flags & 1u) return fail(APHPOIS_ENUMERIC, "system contains NaN or Inf");
  if (method_ == kPcg) {
    if (flags & 2u)
      return fail(APHPOIS_EINVAL,
                  "system is not symmetric (e_xm(c) != e_xp(c-x) etc.); PCG needs a symmetric "
                  "system, use method CONJUGATE");
    if (flags & 4u) return fail(APHPOIS_EINVAL, "zero diagonal coefficient");
    if ((flags & 16u) && (flags & 32u))
      return fail(APHPOIS_EINVAL, "diagonal changes sign; PCG needs a definite system");
    s.neg = (flags & 32u) != 0;
    const bool has_m = (flags & 64u) != 0;
    if (has_m && !s.dev.m) {
      void* m = nullptr;
      APH_TRY(alloc(s, &m, s.n * (fp32_ ? 4 : 8)));
      s.dev.m = m;
    }
    s.has_m = has_m;
    s.dev.has_m = has_m ? 1 : 0;
    s.dev.neg = s.neg ? 1 : 0;
    s.dev.e0_sum = (flags & 128u) ? 0 : 1;
    s.singular = nullspace_ == APHPOIS_NULL_ON || (nullspace_ == APHPOIS_NULL_AUTO && !has_m);
    APH_HIP(launch_convert_pcg(fp32_, d, 0, ly, lz, s.nx, s.ny, s.nz, s.neg ? -1.0 : 1.0,
                               const_cast<void*>(s.dev.kx), const_cast<void*>(s.dev.ky),
                               const_cast<void*>(s.dev.kz), const_cast<void*>(s.dev.S),
                               has_m ? const_cast<void*>(s.dev.m) : nullptr,
                               const_cast<double*>(s.dev.s4[0]), const_cast<double*>(s.dev.b64),
                               0, st_));
    if (pc_) {
      // mass terms of the coarse levels, then coarse operators and D^-1
      for (size_t l = 0; l < s.lev.size(); ++l) {
        MgLevel& L = s.lev[l];
        if (!has_m) {
          L.m = nullptr;
        } else if (l == 0) {
          L.m = s.dev.m;
        } else if (!L.m) {
          void* m = nullptr;
          APH_TRY(alloc(s, &m, L.n * (fp32_ ? 4 : 8)));
          L.m = m;
        }
      }
      APH_HIP(hipMemcpyAsync(s.d_lev, s.lev.data(), sizeof(MgLevel) * s.lev.size(),
                             hipMemcpyHostToDevice, st_));
      APH_HIP(launch_mg_setup(fp32_, s.lev.data(), (int)s.lev.size(), s.dev.kx, s.dev.ky,
                              s.dev.kz, st_));
    }
  } else {
    if (method_ == kJacobi && (flags & 4u))
      return fail(APHPOIS_EINVAL, "zero diagonal coefficient (Jacobi divides by e0)");
    // singular (vanishing row sums): only the max_diff report removes a mean
    s.singular = nullspace_ == APHPOIS_NULL_ON ||
                 (nullspace_ == APHPOIS_NULL_AUTO && (flags & 64u) == 0);
    APH_HIP(launch_convert_f7(d, 0, ly, lz, s.nx, s.ny, s.nz, const_cast<double*>(s.dev.f7[0]),
                              0, st_));
  }
  APH_HIP(hipStreamSynchronize(st_));
  s.has_system = true;
  s.have_x = false;
  dirty_ = true;
  return APHPOIS_OK;
}

int Engine::set_rhs(int slot, const double* e7, bool on_device, long long start, long long ly,
                    long long lz) {
  System* sp = sys(slot);
  if (!sp) return fail(APHPOIS_EINVAL, "no such slot");
  if (!sp->has_system) return fail(APHPOIS_ESTATE, "set_rhs before set_system");
  if (!e7) return fail(APHPOIS_EINVAL, "e7 is NULL");
  System& s = *sp;
  const double* d = nullptr;
  s.bnorm2 = -1;
  APH_TRY(upload_strided(e7, on_device, start, ly, lz, s.nx, s.ny, s.nz, 1, &d));
  if (method_ == kPcg) {
    APH_HIP(launch_convert_rhs(d, 0, ly, lz, s.nx, s.ny, s.nz, -1.0,
                               const_cast<double*>(s.dev.b64), st_));
  } else {
    APH_HIP(launch_convert_rhs(d, 0, ly, lz, s.nx, s.ny, s.nz, 1.0,
                               const_cast<double*>(s.dev.f7[7]), st_));
  }
  APH_HIP(hipStreamSynchronize(st_));
  return APHPOIS_OK;
}

int Engine::set_solution_ref(int slot, const double* xref, long long start, long long ly,
                             long long lz) {
  System* sp = sys(slot);
  if (!sp) return fail(APHPOIS_EINVAL, "no such slot");
  System& s = *sp;
  if (!xref) {
    s.xref = nullptr;  // the storage stays with the system until it is removed
    return APHPOIS_OK;
  }
  if (!s.xref_store) {
    void* p = nullptr;
    APH_TRY(alloc(s, &p, s.n * 8));
    s.xref_store = (double*)p;
  }
  const double* d = nullptr;
  APH_TRY(upload_strided(xref, false, start, ly, lz, s.nx, s.ny, s.nz, 1, &d));
  APH_HIP(launch_gather(d, 0, ly, lz, s.nx, s.ny, s.nz, s.xref_store, 0.0, st_));
  APH_HIP(hipStreamSynchronize(st_));
  s.xref = s.xref_store;
  return APHPOIS_OK;
}

int Engine::prepare(int slot, const double* x0, bool on_device, long long start, long long ly,
                    long long lz, bool keep_x) {
  System* sp = sys(slot);
  if (!sp) return fail(APHPOIS_EINVAL, "no such slot");
  if (!sp->has_system) return fail(APHPOIS_ESTATE, "solve before set_system");
  System& s = *sp;
  const long long n = s.n;
  double* xdst = method_ == kPcg ? s.dev.x64 : s.dev.u[0];
  bool zero_x = false;
  if (x0) {
    const double* d = nullptr;
    APH_TRY(upload_strided(x0, on_device, start, ly, lz, s.nx, s.ny, s.nz, 1, &d));
    APH_HIP(launch_gather(d, 0, ly, lz, s.nx, s.ny, s.nz, xdst, 0.0, st_));
  } else if (keep_x && s.have_x) {
    if (method_ == kPcg) {
      // x64 holds x + mean_pending; fold the pending mean in
      const double mp = h_sc_[slot].mean_pending;
      if (mp != 0.0) APH_HIP(launch_shift(s.dev.x64, n, mp, st_));
    } else if (method_ == kJacobi && h_sc_[slot].cur == 1) {
      APH_HIP(hipMemcpyAsync(s.dev.u[0], s.dev.u[1], n * 8, hipMemcpyDeviceToDevice, st_));
    }
  } else {
    APH_HIP(launch_fill(xdst, n, 8, st_));
    zero_x = true;
  }
  if (method_ == kPcg) {
    // the first CG operator reads p[0] (times beta = 0) and writes p[1]; x^
    // is not filled either: the start's Scalars mark it stale (xh_fold 1)
    const int es = fp32_ ? 4 : 8;
    APH_HIP(launch_fill(s.dev.p[0], n, es, st_));
  }
  // bnorm2 (PCG): sum of b^2, kept until b changes (set_system, set_rhs,
  // perturb_rhs)
  double bnorm2 = 0;
  if (method_ == kPcg) {
    if (s.bnorm2 < 0) {
      APH_HIP(launch_sumsq(s.dev.b64, n, d_red_, d_red_counter_, d_red_ + 1024, st_));
      APH_HIP(hipMemcpyAsync(&bnorm2, d_red_ + 1024, sizeof(double), hipMemcpyDeviceToHost,
                             st_));
      APH_HIP(hipStreamSynchronize(st_));
      s.bnorm2 = bnorm2;
    }
    bnorm2 = s.bnorm2;
  }
  Scalars sc{};
  sc.xh_fold = 1;  // x^ holds no value yet
  sc.tol = tol_;
  sc.delta = delta_;
  sc.h3 = s.h3;
  sc.bnorm2 = bnorm2;
  sc.gamma = 1.0;
  sc.iter = 0;
  sc.miniter = miniter_;
  sc.maxiter = maxiter_;
  sc.done = 0;
  sc.ru = method_ == kPcg ? 1 : 0;
  sc.resmode = resmode_;
  sc.singular = s.singular ? 1 : 0;
  sc.first = 1;
  sc.cur = 0;
  sc.method = method_;
  sc.pc = pc_ ? 1 : 0;
  // x64 and x^ zero-filled above: the start's residual is b itself
  sc.zero_start = method_ == kPcg && zero_x ? 1 : 0;
  h_sc_[slot] = sc;
  APH_HIP(hipMemcpyAsync(d_sc_ + slot, h_sc_ + slot, sizeof(Scalars), hipMemcpyHostToDevice,
                         st_));
  if (!s.active) dirty_ = true;
  s.active = true;
  s.done_reported = false;
  need_start_ = true;
  return APHPOIS_OK;
}

int Engine::build_tables() {
  launch_slots_.clear();
  bool all4 = true, all2 = true;
  for (int i = 0; i < (int)slots_.size(); ++i) {
    System* s = slots_[i].get();
    if (s && s->active) {
      launch_slots_.push_back(i);
      all4 = all4 && s->nx % 4 == 0;
      all2 = all2 && s->nx % 2 == 0;
    }
  }
  if (method_ == kPcg) {
    vx_ = fp32_ ? (all4 ? 4 : (all2 ? 2 : 1)) : (all2 ? 2 : 1);
    const int tv = tuning_.vx;
    if (tv == 1 || (tv == 2 && all2) || (tv == 4 && all4 && fp32_)) vx_ = tv;
  } else {
    vx_ = 1;
  }
  // the PCG FP64 residual: 2 cells per thread when every row has an even
  // length (the reference modes' residual stays at 1)
  const int vx_ru = method_ == kPcg && all2 ? 2 : 1;
  const int ns = (int)launch_slots_.size();
  std::vector<Rec> rst, rsm, rru;
  std::vector<int> bst, bsm, bru;
  std::vector<SysDev> sds;
  // multigrid: per level l, tilings of level l (lf) and l+1 (lc)
  std::vector<std::vector<Rec>> rlf(kMaxLevels), rlc(kMaxLevels);
  std::vector<std::vector<int>> blf(kMaxLevels), blc(kMaxLevels);
  int nlev_max = 0;
  long long coarse_max = 0;  // LDS bytes of the bottom kernel
  int bottom_at[kMaxLevels] = {};
  int deep = 0;
  // x vector width of each level's sweeps: the widest that divides the level's
  // nx of every system that smooths on that level
  int vxl[kMaxLevels];
  for (int l = 0; l < kMaxLevels; ++l) vxl[l] = fp32_ ? 4 : 2;
  if (pc_) {
    for (int i : launch_slots_) {
      const System& s = *slots_[i];
      for (int l = 0; l + 1 < (int)s.lev.size(); ++l) {
        while (vxl[l] > 1 && s.lev[l].nx % vxl[l] != 0) vxl[l] /= 2;
      }
    }
  }
  // post-smoothing per level: fused for the systems whose level says so
  int post_fused[kMaxLevels] = {}, post_plain[kMaxLevels] = {}, post_tile[kMaxLevels] = {};
  std::vector<std::vector<Rec>> rlp(kMaxLevels), rlt(kMaxLevels);
  std::vector<std::vector<int>> blp(kMaxLevels), blt(kMaxLevels);
  for (int k = 0; k < ns; ++k) {
    System& s = *slots_[launch_slots_[k]];
    // per-system tile (autotuner) over the engine-wide override
    Tuning tu = tuning_;
    if (s.tune_lx > 0) tu.lx = s.tune_lx;
    if (s.tune_bz > 0) tu.bz = s.tune_bz;
    Rec a, b, c;
    system_tiling(method_, s.nx, s.ny, s.nz, vx_, tu, &a, &b, &c, vx_ru, fp32_ ? 0 : 1);
    a.sys = b.sys = c.sys = k;
    a.blk0 = (int)bst.size();
    b.blk0 = (int)bsm.size();
    c.blk0 = (int)bru.size();
    Tuning none;
    for (int i = 0; i < a.nblk; ++i) bst.push_back((int)rst.size());
    for (int i = 0; i < b.nblk; ++i) bsm.push_back((int)rsm.size());
    for (int i = 0; i < c.nblk; ++i) bru.push_back((int)rru.size());
    rst.push_back(a);
    rsm.push_back(b);
    rru.push_back(c);
    int nblk_pc = 0;
    if (pc_) {
      const int nl = (int)s.lev.size();
      nlev_max = std::max(nlev_max, nl);
      // LDS of the bottom kernel (same rule as the kernel: everything
      // resident if it fits, else z and a second buffer per level)
      const int es = fp32_ ? 4 : 8;
      long long zbytes = 0, fbytes = 0;
      for (int l = s.dev.mg_lb; l < nl; ++l) {
        zbytes += bottom_z_bytes(s.lev[l].n, es);
        fbytes += bottom_full_bytes(s.lev[l].n, es, s.dev.has_m);
      }
      coarse_max = std::max(coarse_max, fbytes <= kBottomLdsMax ? fbytes : zbytes);
      bottom_at[s.dev.mg_lb] = 1;
      deep = std::max(deep, s.dev.mg_lb);
      for (int l = 0; l < s.dev.mg_lb; ++l) {
        const MgLevel& F = s.lev[l];
        const MgLevel& C = s.lev[l + 1];
        // large levels: long z marches as above; small levels (latency
        // bound): short marches over more workgroups
        const long long ftgt = F.n >= (1LL << 21) ? std::min(1024LL, cdiv(F.n, 2048))
                                                  : std::max(1LL, cdiv(F.n, 512));
        // level 0 follows the system's stencil tuning (autotuner / set_tile)
        Rec f = make_stencil_rec(F.nx, F.ny, F.nz, vxl[l], ftgt, l == 0 ? tu : none);
        // the fused restriction sums plane pairs inside one workgroup: even
        // plane counts per tile (F.nz is even on every coarsened level)
        if (f.bz & 1) {
          f.bz = std::min(f.bz + 1, F.nz);
          f.tiles_z = (int)cdiv(F.nz, f.bz);
          f.nblk = f.tiles_x * f.tiles_y * f.tiles_z;
        }
        f.sys = k;
        f.blk0 = (int)blf[l].size();
        for (int i = 0; i < f.nblk; ++i) blf[l].push_back((int)rlf[l].size());
        rlf[l].push_back(f);
        Rec g = make_stencil_rec(C.nx, C.ny, C.nz, 1,
                                 std::min(3072LL, std::max(1LL, cdiv(C.n, 2048))), none);
        g.sys = k;
        g.blk0 = (int)blc[l].size();
        for (int i = 0; i < g.nblk; ++i) blc[l].push_back((int)rlc[l].size());
        rlc[l].push_back(g);
        if (l == 0) nblk_pc = f.nblk;
        if (F.post == 1) {
          post_fused[l] = 1;
          Rec p{};
          make_post_rec(F.nx, F.ny, F.nz, fp32_ ? 4 : 2, ncu_, &p);
          p.sys = k;
          p.blk0 = (int)blp[l].size();
          for (int i = 0; i < p.nblk; ++i) blp[l].push_back((int)rlp[l].size());
          rlp[l].push_back(p);
          if (l == 0) nblk_pc = p.nblk;
        } else if (F.post == 2) {
          post_tile[l] = 1;
          Rec t{};
          t.tiles_x = (int)cdiv(F.nx, kTile);
          t.tiles_y = (int)cdiv(F.ny, kTile);
          t.tiles_z = (int)cdiv(F.nz, kTile);
          t.lx = kTile;
          t.bz = kTile;
          t.nblk = t.tiles_x * t.tiles_y * t.tiles_z;
          t.sys = k;
          t.blk0 = (int)blt[l].size();
          for (int i = 0; i < t.nblk; ++i) blt[l].push_back((int)rlt[l].size());
          rlt[l].push_back(t);
        } else {
          post_plain[l] = 1;
        }
      }
    }
    s.dev.nblk_pc = nblk_pc;
    // partials: 3 values per block of the largest reducing launch of this system
    const int nmax = std::max(std::max(a.nblk, nblk_pc), std::max(b.nblk, c.nblk));
    const long long need = 3LL * nmax * 8;
    if (need > s.partials_cap) {
      if (s.partials) APH_HIP(hipFree(s.partials));
      s.partials = nullptr;
      s.partials_cap = 0;
      APH_HIP(hipMalloc(&s.partials, need));
      s.partials_cap = need;
    }
    s.dev.partials = (double*)s.partials;
    s.dev.nblk_st = a.nblk;
    s.dev.nblk_sm = b.nblk;
    s.dev.nblk_ru = c.nblk;
    sds.push_back(s.dev);
  }
  // device tables in one buffer each
  long long nrec = (long long)rst.size() + rsm.size() + rru.size();
  long long nb2r = (long long)bst.size() + bsm.size() + bru.size();
  for (int l = 0; l < kMaxLevels; ++l) {
    nrec += (long long)rlf[l].size() + rlc[l].size() + rlp[l].size() + rlt[l].size();
    nb2r += (long long)blf[l].size() + blc[l].size() + blp[l].size() + blt[l].size();
  }
  if (nrec > rec_cap_) {
    if (d_rec_) APH_HIP(hipFree(d_rec_));
    APH_HIP(hipMalloc(&d_rec_, sizeof(Rec) * std::max(1LL, nrec)));
    rec_cap_ = nrec;
  }
  if (nb2r > b2r_cap_) {
    if (d_b2r_) APH_HIP(hipFree(d_b2r_));
    APH_HIP(hipMalloc(&d_b2r_, sizeof(int) * std::max(1LL, nb2r)));
    b2r_cap_ = nb2r;
  }
  if (d_sds_) APH_HIP(hipFree(d_sds_));
  d_sds_ = nullptr;
  APH_HIP(hipMalloc(&d_sds_, sizeof(SysDev) * std::max(1, ns)));
  std::vector<Rec> recs;
  recs.insert(recs.end(), rst.begin(), rst.end());
  recs.insert(recs.end(), rsm.begin(), rsm.end());
  recs.insert(recs.end(), rru.begin(), rru.end());
  std::vector<int> b2r;
  b2r.insert(b2r.end(), bst.begin(), bst.end());
  for (int v : bsm) b2r.push_back(v);
  for (int v : bru) b2r.push_back(v);
  size_t rec_lf_ofs[kMaxLevels], rec_lc_ofs[kMaxLevels], b2r_lf_ofs[kMaxLevels],
      b2r_lc_ofs[kMaxLevels], rec_lp_ofs[kMaxLevels], b2r_lp_ofs[kMaxLevels],
      rec_lt_ofs[kMaxLevels], b2r_lt_ofs[kMaxLevels];
  for (int l = 0; l < kMaxLevels; ++l) {
    rec_lf_ofs[l] = recs.size();
    recs.insert(recs.end(), rlf[l].begin(), rlf[l].end());
    rec_lc_ofs[l] = recs.size();
    recs.insert(recs.end(), rlc[l].begin(), rlc[l].end());
    rec_lp_ofs[l] = recs.size();
    recs.insert(recs.end(), rlp[l].begin(), rlp[l].end());
    rec_lt_ofs[l] = recs.size();
    recs.insert(recs.end(), rlt[l].begin(), rlt[l].end());
    b2r_lf_ofs[l] = b2r.size();
    b2r.insert(b2r.end(), blf[l].begin(), blf[l].end());
    b2r_lc_ofs[l] = b2r.size();
    b2r.insert(b2r.end(), blc[l].begin(), blc[l].end());
    b2r_lp_ofs[l] = b2r.size();
    b2r.insert(b2r.end(), blp[l].begin(), blp[l].end());
    b2r_lt_ofs[l] = b2r.size();
    b2r.insert(b2r.end(), blt[l].begin(), blt[l].end());
  }
  if (ns > 0) {
    APH_HIP(hipMemcpy(d_sds_, sds.data(), sizeof(SysDev) * ns, hipMemcpyHostToDevice));
    APH_HIP(hipMemcpy(d_rec_, recs.data(), sizeof(Rec) * nrec, hipMemcpyHostToDevice));
    APH_HIP(hipMemcpy(d_b2r_, b2r.data(), sizeof(int) * nb2r, hipMemcpyHostToDevice));
  }
  L_.sds = d_sds_;
  L_.rec_st = d_rec_;
  L_.rec_sm = d_rec_ + rst.size();
  L_.rec_ru = d_rec_ + rst.size() + rsm.size();
  L_.b2r_st = d_b2r_;
  L_.b2r_sm = d_b2r_ + bst.size();
  L_.b2r_ru = d_b2r_ + bst.size() + bsm.size();
  L_.nst = (int)bst.size();
  L_.nsm = (int)bsm.size();
  L_.nru = (int)bru.size();
  L_.vx_ru = vx_ru;
  L_.ns = ns;
  L_.pc = pc_ ? 1 : 0;
  L_.nlev_max = nlev_max;
  L_.mg_w = 0;
  for (int i : launch_slots_) L_.mg_w = std::max(L_.mg_w, slots_[i]->dev.mg_w);
  for (int l = 0; l < kMaxLevels; ++l) L_.vx_l[l] = vxl[l];
  L_.coarse_sweeps = kCoarseSweeps;
  L_.coarse_lds = (int)coarse_max;
  for (int l = 0; l < kMaxLevels; ++l) L_.bottom_at[l] = bottom_at[l];
  L_.deep = deep;
  for (int l = 0; l < kMaxLevels; ++l) {
    L_.rec_lf[l] = d_rec_ + rec_lf_ofs[l];
    L_.rec_lc[l] = d_rec_ + rec_lc_ofs[l];
    L_.b2r_lf[l] = d_b2r_ + b2r_lf_ofs[l];
    L_.b2r_lc[l] = d_b2r_ + b2r_lc_ofs[l];
    L_.nlf[l] = (int)blf[l].size();
    L_.nlc[l] = (int)blc[l].size();
    L_.post_fused[l] = post_fused[l];
    L_.post_plain[l] = post_plain[l];
    L_.vx_p[l] = fp32_ ? 4 : 2;
    L_.rec_lp[l] = d_rec_ + rec_lp_ofs[l];
    L_.b2r_lp[l] = d_b2r_ + b2r_lp_ofs[l];
    L_.nlp[l] = (int)blp[l].size();
    L_.post_tile[l] = post_tile[l];
    L_.rec_lt[l] = d_rec_ + rec_lt_ofs[l];
    L_.b2r_lt[l] = d_b2r_ + b2r_lt_ofs[l];
    L_.nlt[l] = (int)blt[l].size();
  }
  dirty_ = false;
  return APHPOIS_OK;
}

int Engine::capture() {
  if (gexec_) {
    APH_HIP(hipGraphExecDestroy(gexec_));
    gexec_ = nullptr;
  }
  if (graph_) {
    APH_HIP(hipGraphDestroy(graph_));
    graph_ = nullptr;
  }
  if (below_exec_) {
    APH_HIP(hipGraphExecDestroy(below_exec_));
    below_exec_ = nullptr;
  }
  if (below_graph_) {
    APH_HIP(hipGraphDestroy(below_graph_));
    below_graph_ = nullptr;
  }
  const int nk = iter_kernels(method_, pc_ ? 1 : 0);
  const int ng = graph_kernels(method_, pc_ ? 1 : 0);  // of them in the graph
  // launches per iteration (the multigrid cycle counts as one kernel in nk)
  const int nl = ng + (pc_ ? vcycle_launches(L_) - 1 : 0);
  int k = check_every_;
  if (k <= 0) {
    // about 4 ms of work per replay (iterations past convergence cost a
    // launch each), at most ~2000 graph nodes
    long long ntot = 0;
    for (int i : launch_slots_) ntot += slots_[i]->n;
    const double bpc = method_ == kPcg ? (fp32_ ? 60.0 : 120.0) * (pc_ ? 2.5 : 1.0) : 150.0;
    const double t_iter = (double)ntot * bpc / 4e12 + nl * 6e-6;
    k = (int)std::lround(4e-3 / t_iter);
    k = std::max(2, std::min({64, k, std::max(2, 2000 / nl)}));
  }
  if (method_ == kJacobi && (k & 1)) ++k;
  if (profile_) {
    const size_t ne = (size_t)k * nk + 1;
    while (pev_.size() < ne) {
      // timing only (read after the stream is synchronized): no system-scope
      // fence, whose L2 writeback would slow the kernel after each event
      hipEvent_t e;
      APH_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
      pev_.push_back(e);
    }
  }
  APH_HIP(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
  for (int i = 0; i < k; ++i) {
    for (int j = 0; j < ng; ++j) {
      const hipError_t e = launch_kernel(method_, fp32_, vx_, L_, st_, j);
      if (e != hipSuccess) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(st_, &g);
        if (g) (void)hipGraphDestroy(g);
        return hip(e, "launch (capture)");
      }
    }
  }
  APH_HIP(hipStreamEndCapture(st_, &graph_));
  APH_HIP(hipGraphInstantiate(&gexec_, graph_, nullptr, nullptr, 0));
  graph_k_ = k;
  return APHPOIS_OK;
}

// After a replay of a graph without the reliable-update kernels
// (graph_kernels): systems whose update triggered one have idled since; the
// reliable-update kernels of the start (launch_start, which skip the other
// systems) apply it, and the scalars are read again for its stop test.
int Engine::run_pending_ru() {
  if (graph_kernels(method_, pc_ ? 1 : 0) == iter_kernels(method_, pc_ ? 1 : 0)) {
    return APHPOIS_OK;
  }
  bool pending = false;
  for (int i : launch_slots_) pending = pending || (h_sc_[i].ru && !h_sc_[i].done);
  if (!pending) return APHPOIS_OK;
  APH_HIP(launch_start(method_, fp32_, vx_, L_, st_));
  return pull_scalars();
}

// Copies the scalars of every system in the launch tables (one contiguous
// range of slots) to the pinned mirror and waits for the stream.
int Engine::pull_scalars() {
  if (launch_slots_.empty()) return APHPOIS_OK;
  const int lo = launch_slots_.front(), hi = launch_slots_.back();
  APH_HIP(hipMemcpyAsync(h_sc_ + lo, d_sc_ + lo, sizeof(Scalars) * (hi - lo + 1),
                         hipMemcpyDeviceToHost, st_));
  APH_HIP(hipStreamSynchronize(st_));
  return APHPOIS_OK;
}

void Engine::fill_info(int slot, double seconds) {
  System& s = *slots_[slot];
  const Scalars& c = h_sc_[slot];
  aphpois_info& inf = s.info;
  std::memset(&inf, 0, sizeof(inf));
  inf.iter = c.iter;
  inf.residual = c.residual;
  inf.res_l2 = c.res_l2;
  inf.res_max = c.res_max;
  inf.true_relres = c.relres;
  inf.mean = c.mean_pending;
  inf.converged = c.converged;
  inf.refinements = c.nru;
  inf.status = c.status;
  inf.seconds = seconds;
  inf.max_diff = -1.0;
}

int Engine::diff_info(int slot) {
  System& s = *slots_[slot];
  if (!s.xref) return APHPOIS_OK;
  const double* x;
  double sub = 0;
  if (method_ == kPcg) {
    x = s.dev.x64;
    sub = h_sc_[slot].mean_pending;
  } else if (method_ == kJacobi) {
    x = s.dev.u[h_sc_[slot].cur];
  } else {
    x = s.dev.u[0];
  }
  double out[2] = {0, 0};
  APH_HIP(launch_max_diff(x, s.xref, s.n, sub, s.singular ? 1 : 0, d_red_, d_red_ + 3072, st_));
  APH_HIP(hipMemcpyAsync(out, d_red_ + 3072, sizeof(out), hipMemcpyDeviceToHost, st_));
  APH_HIP(hipStreamSynchronize(st_));
  s.info.max_diff = out[1];
  return APHPOIS_OK;
}

int Engine::perturb_rhs(int slot, double amp, unsigned long long key) {
  System* sp = sys(slot);
  if (!sp) return fail(APHPOIS_EINVAL, "no such slot");
  if (method_ != kPcg || !sp->has_system) return fail(APHPOIS_ESTATE, "no PCG system");
  sp->bnorm2 = -1;
  APH_HIP(launch_perturb(const_cast<double*>(sp->dev.b64), sp->n, key, amp, d_red_, st_));
  return APHPOIS_OK;
}

int Engine::precond_apply(int slot, const double* r, double* z) {
  System* sp = sys(slot);
  if (!sp) return fail(APHPOIS_EINVAL, "no such slot");
  if (method_ != kPcg || !pc_) return fail(APHPOIS_EINVAL, "needs method PCG with multigrid");
  if (!sp->has_system) return fail(APHPOIS_ESTATE, "no system set");
  System& s = *sp;
  const long long n = s.n;
  const int es = fp32_ ? 4 : 8;
  // only this system in the launch tables
  std::vector<char> was_active(slots_.size(), 0);
  for (size_t i = 0; i < slots_.size(); ++i) {
    if (slots_[i]) {
      was_active[i] = slots_[i]->active;
      slots_[i]->active = (int)i == slot;
    }
  }
  int rc = build_tables();
  auto restore = [&]() {
    for (size_t i = 0; i < slots_.size(); ++i)
      if (slots_[i]) slots_[i]->active = was_active[i];
    dirty_ = true;
  };
  if (rc != APHPOIS_OK) {
    restore();
    return rc;
  }
  const double* d = nullptr;
  rc = upload_strided(r, false, 0, s.nx, (long long)s.nx * s.ny, s.nx, s.ny, s.nz, 1, &d);
  Scalars sc{};
  sc.pc = 1;
  sc.first = 1;
  sc.method = method_;
  hipError_t e = hipSuccess;
  if (rc == APHPOIS_OK) {
    e = launch_cast(d, 8, const_cast<void*>(s.lev[0].r), es, n, st_);
    if (e == hipSuccess) e = hipStreamSynchronize(st_);
    if (e == hipSuccess) e = hipMemcpy(d_sc_ + slot, &sc, sizeof(Scalars), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_kernel(method_, fp32_, vx_, L_, st_, 8);
    // z of level 0 back through the staging buffer (free again after the upload)
    if (e == hipSuccess) e = launch_cast(s.lev[0].z[1], es, staging_, 8, n, st_);
    if (e == hipSuccess) e = hipMemcpyAsync(z, staging_, n * 8, hipMemcpyDeviceToHost, st_);
    if (e == hipSuccess) e = hipStreamSynchronize(st_);
    // the device scalars of the slot back to their last pulled state
    if (e == hipSuccess)
      e = hipMemcpy(d_sc_ + slot, h_sc_ + slot, sizeof(Scalars), hipMemcpyHostToDevice);
  }
  restore();
  if (rc != APHPOIS_OK) return rc;
  return hip(e, "precond_apply");
}

int Engine::step(int max_replays, std::vector<int>* finished) {
  finished->clear();
  if (method_ != kPcg) return fail(APHPOIS_EINVAL, "step() needs method PCG");
  bool any = false;
  for (auto& s : slots_) any = any || (s && s->active);
  if (!any) return APHPOIS_OK;
  if (dirty_) {
    std::unique_lock<std::mutex> lk;
    if (api_mu_) lk = std::unique_lock<std::mutex>(*api_mu_);
    APH_TRY(build_tables());
    APH_TRY(capture());
  }
  if (need_start_) {
    // reliable-update kernels: only systems prepared since the last call
    // (ru = 1) take part, running systems are at an iteration boundary
    APH_HIP(launch_start(method_, fp32_, vx_, L_, st_));
    need_start_ = false;
  }
  for (int r = 0; r < max_replays && finished->empty(); ++r) {
    APH_HIP(hipGraphLaunch(gexec_, st_));
    APH_TRY(pull_scalars());
    APH_TRY(run_pending_ru());
    bool alldone = true;
    for (int i : launch_slots_) {
      System& s = *slots_[i];
      if (h_sc_[i].done && !s.done_reported) {
        s.done_reported = true;
        s.have_x = true;
        fill_info(i, 0.0);
        APH_TRY(diff_info(i));
        finished->push_back(i);
      }
      alldone = alldone && h_sc_[i].done;
    }
    if (alldone) break;
  }
  return APHPOIS_OK;
}

int Engine::run() {
  bool any = false;
  for (auto& s : slots_) any = any || (s && s->active);
  if (!any) return APHPOIS_OK;
  if (dirty_) {
    APH_TRY(build_tables());
    APH_TRY(capture());
  }
  APH_HIP(hipEventRecord(ev0_, st_));
  APH_HIP(launch_start(method_, fp32_, vx_, L_, st_));
  // a replay may end with a system idle until its reliable update (graph
  // without those kernels); every replay still advances each system by at
  // least one iteration or ends it
  const int ng = graph_kernels(method_, pc_ ? 1 : 0);
  const int maxloops = (std::max(maxiter_, miniter_) + 1) / std::max(1, graph_k_) + 4 +
                       (ng < iter_kernels(method_, pc_ ? 1 : 0) ? maxiter_ + 1 : 0);
  int loops = 0;
  bool alldone = false;
  int iters_before = 0;
  while (!alldone) {
    if (profile_) {
      // same kernels as the graph, launched one by one with an event before
      // each kernel (HIP does not record events captured into a graph); the
      // cycle below level 0 is one replay of its own graph, so that the
      // event-timed solve launches only the level-0 and CG kernels singly
      const int nk = iter_kernels(method_, pc_ ? 1 : 0);
      const int nv = pc_ ? vcycle_launches(L_) + 1 : 0;
      while ((int)pev_mg_.size() < graph_k_ * nv) {
        hipEvent_t e;  // timing only, as pev_
        APH_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
        pev_mg_.push_back(e);
      }
      mg_kinds_.resize(nv);
      LaunchSet Lp = L_;
      if (pc_) {
        if (!below_exec_) {
          APH_HIP(hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal));
          const hipError_t e = capture_below0(fp32_, L_, st_);
          hipGraph_t g = nullptr;
          const hipError_t e2 = hipStreamEndCapture(st_, &g);
          if (e != hipSuccess || e2 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return hip(e != hipSuccess ? e : e2, "capture (cycle below level 0)");
          }
          below_graph_ = g;
          APH_HIP(hipGraphInstantiate(&below_exec_, below_graph_, nullptr, nullptr, 0));
        }
        Lp.below0 = below_exec_;
      }
      for (int sl = 0; sl < graph_k_; ++sl) {
        for (int j = 0; j < nk; ++j) {
          if (prof_timed(j)) APH_HIP(hipEventRecord(pev_[(size_t)sl * nk + j], st_));
          if (j == 8 && pc_) {
            int nl = 0;
            APH_HIP(launch_vcycle_profiled(fp32_, Lp, st_, &pev_mg_[(size_t)sl * nv],
                                           mg_kinds_.data(), &nl));
            prof_nl_ = nl;
          } else {
            APH_HIP(launch_kernel(method_, fp32_, vx_, L_, st_, j));
          }
        }
      }
      APH_HIP(hipEventRecord(pev_[(size_t)graph_k_ * nk], st_));
    } else {
      APH_HIP(hipGraphLaunch(gexec_, st_));
    }
    ++loops;
    APH_TRY(pull_scalars());
    if (!profile_) APH_TRY(run_pending_ru());
    alldone = true;
    int iters_after = 0;
    for (int i : launch_slots_) {
      alldone = alldone && h_sc_[i].done;
      iters_after = std::max(iters_after, h_sc_[i].iter);
    }
    if (profile_) {
      // slots that ran a full iteration for at least one system
      const int nk = iter_kernels(method_, pc_ ? 1 : 0);
      const int ran = std::min(graph_k_, iters_after - iters_before);
      for (int sl = 0; sl < ran; ++sl) {
        for (int j = 0; j < nk; ++j) {
          if (!prof_timed(j)) continue;
          int j2 = j + 1;  // the next event: the next timed kernel's, or the slot's end
          while (j2 < nk && !prof_timed(j2)) ++j2;
          float ms = 0;
          APH_HIP(hipEventElapsedTime(&ms, pev_[(size_t)sl * nk + j],
                                      pev_[(size_t)sl * nk + j2]));
          prof_ms_[j] += ms;
          prof_n_[j] += 1;
        }
        if (pc_) {
          // cycle launches by kind: [9] level-0 sweep (per launch, all
          // three), [10] level-0 restriction, [11] coarser levels (per
          // cycle), [12] bottom-levels kernel, [13] [14] [15] level-0
          // pre-sweep, first and second post-sweep
          const int nv = vcycle_launches(L_) + 1;  // event stride per cycle
          double coarse = 0;
          for (int i = 0; i < prof_nl_; ++i) {
            float ms = 0;
            APH_HIP(hipEventElapsedTime(&ms, pev_mg_[(size_t)sl * nv + i],
                                        pev_mg_[(size_t)sl * nv + i + 1]));
            const int kd = mg_kinds_[i];
            int slot2 = -1, slot1 = -1;
            if (kd == 0 || kd == 5 || kd == 6) {
              slot1 = 9;
              slot2 = kd == 0 ? 13 : (kd == 5 ? 14 : 15);
            } else if (kd == 1) {
              slot1 = 10;
            } else if (kd == 3) {
              slot1 = 12;
            } else if (kd == 2) {
              coarse += ms;
            }
            if (slot1 >= 0) {
              prof_ms_[slot1] += ms;
              prof_n_[slot1] += 1;
            }
            if (slot2 >= 0) {
              prof_ms_[slot2] += ms;
              prof_n_[slot2] += 1;
            }
          }
          prof_ms_[11] += coarse;
          prof_n_[11] += 1;
        }
      }
    }
    iters_before = iters_after;
    if (loops > maxloops + 2 && !alldone) {
      // every system hits maxiter within maxloops replays; reaching here means
      // a flag was never set
      return fail(APHPOIS_EDEVICE, "solver did not terminate within maxiter");
    }
  }
  APH_HIP(hipEventRecord(ev1_, st_));
  if (method_ != kPcg) {
    APH_HIP(launch_true_res_f7(L_, st_));
  }
  APH_TRY(pull_scalars());
  float ms = 0;
  APH_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
  last_run_ms_ = ms;
  last_run_iters_ = 0;
  for (int i : launch_slots_) {
    System& s = *slots_[i];
    if (s.done_reported) continue;  // finished in an earlier call, idle since
    fill_info(i, ms * 1e-3);
    APH_TRY(diff_info(i));
    last_run_iters_ = std::max(last_run_iters_, h_sc_[i].iter);
    s.have_x = true;
    s.done_reported = true;
  }
  // systems stay in the launch tables (idle once done); preparing them again
  // reuses the captured graph
  need_start_ = false;
  return APHPOIS_OK;
}

int Engine::fetch(int slot, double* x, bool on_device, long long start, long long ly,
                  long long lz, aphpois_info* info) {
  System* sp = sys(slot);
  if (!sp) return fail(APHPOIS_EINVAL, "no such slot");
  System& s = *sp;
  if (info) *info = s.info;
  if (!x) return APHPOIS_OK;
  if (!s.have_x) return fail(APHPOIS_ESTATE, "no solution yet");
  const double* src;
  double sub = 0;
  if (method_ == kPcg) {
    src = s.dev.x64;
    sub = h_sc_[slot].mean_pending;
  } else if (method_ == kJacobi) {
    src = s.dev.u[h_sc_[slot].cur];
  } else {
    src = s.dev.u[0];
  }
  const bool dense = start == 0 && ly == s.nx && lz == (long long)s.nx * s.ny;
  if (on_device) {
    APH_HIP(launch_scatter(src, s.nx, s.ny, s.nz, x, start, ly, lz, sub, st_));
    APH_HIP(hipStreamSynchronize(st_));
    return APHPOIS_OK;
  }
  // dense result in the staging buffer, then host copy
  APH_TRY(ensure_staging(s.n * 8));
  double* st = (double*)staging_;
  APH_HIP(launch_scatter(src, s.nx, s.ny, s.nz, st, 0, s.nx, (long long)s.nx * s.ny, sub, st_));
  if (dense) {
    APH_HIP(hipMemcpyAsync(x, st, s.n * 8, hipMemcpyDeviceToHost, st_));
    APH_HIP(hipStreamSynchronize(st_));
  } else {
    std::vector<double> tmp(s.n);
    APH_HIP(hipMemcpyAsync(tmp.data(), st, s.n * 8, hipMemcpyDeviceToHost, st_));
    APH_HIP(hipStreamSynchronize(st_));
    for (int z = 0; z < s.nz; ++z)
      for (int y = 0; y < s.ny; ++y)
        std::memcpy(x + start + z * lz + y * ly, tmp.data() + ((long long)z * s.ny + y) * s.nx,
                    sizeof(double) * s.nx);
  }
  return APHPOIS_OK;
}

}  // namespace aph
