// C ABI of aphpois (include/aphpois.h) over the Engine.
#include <cstring>
#include <new>
#include <string>

#include "engine.h"

namespace {

thread_local std::string g_err;

int set_err(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

}  // namespace

namespace aph {
int set_api_error(int code, const std::string& msg) { return set_err(code, msg); }
}  // namespace aph

struct aphpois {
  aph::Engine eng;
  aphpois_desc desc{};
  int slot = -1;
};

struct aphpois_batch {
  aph::Engine eng;
  aphpois_batch_desc desc{};
};

#define APH_CHECK_H(h) \
  if (!(h)) return set_err(APHPOIS_EINVAL, "NULL handle")
#define APH_RET(obj, expr)                                  \
  do {                                                      \
    const int rc__ = (expr);                                \
    if (rc__ != APHPOIS_OK) return set_err(rc__, (obj).eng.err()); \
    return APHPOIS_OK;                                      \
  } while (0)

extern "C" {

const char* aphpois_last_error(void) { return g_err.c_str(); }

void aphpois_desc_default(aphpois_desc* d) {
  if (!d) return;
  std::memset(d, 0, sizeof(*d));
  d->nx = d->ny = d->nz = 0;
  d->h3 = 1.0;
  d->method = APHPOIS_CONJUGATE;
  d->precision = APHPOIS_FP64;
  d->resmode = APHPOIS_RES_L2;
  d->tol = 0;
  d->miniter = 1;
  d->maxiter = 100;
  d->nullspace = APHPOIS_NULL_AUTO;
  d->delta = 0;
  d->check_every = 0;
  d->device = 0;
  d->mem_budget = 0;
  d->mg_w = APHPOIS_MG_W_AUTO;
}

int aphpois_create(const aphpois_desc* desc, aphpois** out) {
  if (!desc || !out) return set_err(APHPOIS_EINVAL, "NULL argument");
  *out = nullptr;
  aphpois* h = new (std::nothrow) aphpois();
  if (!h) return set_err(APHPOIS_ENOMEM, "host allocation failed");
  h->desc = *desc;
  int rc = h->eng.init(desc->device, desc->method, desc->precision, desc->resmode, desc->tol,
                       desc->miniter, desc->maxiter, desc->nullspace, desc->delta,
                       desc->check_every, desc->mem_budget, 1, desc->precond, desc->mg_w);
  if (rc == APHPOIS_OK) {
    rc = h->eng.add_system(desc->nx, desc->ny, desc->nz, desc->h3, &h->slot);
  }
  if (rc != APHPOIS_OK) {
    set_err(rc, h->eng.err());
    delete h;
    return rc;
  }
  *out = h;
  return APHPOIS_OK;
}

int aphpois_set_system(aphpois* h, const double* expr, long long start, long long lead_y,
                       long long lead_z) {
  APH_CHECK_H(h);
  APH_RET(*h, h->eng.set_system(h->slot, expr, false, start, lead_y, lead_z));
}

int aphpois_set_system_device(aphpois* h, const double* expr_dev) {
  APH_CHECK_H(h);
  const auto& d = h->desc;
  APH_RET(*h, h->eng.set_system(h->slot, expr_dev, true, 0, d.nx, (long long)d.nx * d.ny));
}

int aphpois_set_rhs(aphpois* h, const double* e7, long long start, long long lead_y,
                    long long lead_z) {
  APH_CHECK_H(h);
  APH_RET(*h, h->eng.set_rhs(h->slot, e7, false, start, lead_y, lead_z));
}

static int solve_impl(aphpois* h, const double* x0, double* x, bool dev, long long start,
                      long long ly, long long lz, aphpois_info* info) {
  if (!x) return set_err(APHPOIS_EINVAL, "x is NULL");
  int rc = h->eng.prepare(h->slot, x0, dev, start, ly, lz, false);
  if (rc == APHPOIS_OK) rc = h->eng.run();
  if (rc == APHPOIS_OK) rc = h->eng.fetch(h->slot, x, dev, start, ly, lz, info);
  if (rc != APHPOIS_OK) return set_err(rc, h->eng.err());
  return APHPOIS_OK;
}

int aphpois_solve(aphpois* h, const double* x0, double* x, long long start, long long lead_y,
                  long long lead_z, aphpois_info* info) {
  APH_CHECK_H(h);
  return solve_impl(h, x0, x, false, start, lead_y, lead_z, info);
}

int aphpois_solve_device(aphpois* h, const double* x0_dev, double* x_dev, aphpois_info* info) {
  APH_CHECK_H(h);
  const auto& d = h->desc;
  return solve_impl(h, x0_dev, x_dev, true, 0, d.nx, (long long)d.nx * d.ny, info);
}

int aphpois_set_solution_ref(aphpois* h, const double* xref, long long start,
                             long long lead_y, long long lead_z) {
  APH_CHECK_H(h);
  APH_RET(*h, h->eng.set_solution_ref(h->slot, xref, start, lead_y, lead_z));
}

int aphpois_set_conf(aphpois* h, double tol, int miniter, int maxiter) {
  APH_CHECK_H(h);
  if (maxiter < 0) return set_err(APHPOIS_EINVAL, "maxiter < 0");
  h->desc.tol = tol;
  h->desc.miniter = miniter;
  h->desc.maxiter = maxiter;
  h->eng.set_conf(tol, miniter, maxiter);
  return APHPOIS_OK;
}

long long aphpois_device_bytes(const aphpois* h) { return h ? h->eng.bytes() : 0; }

int aphpois_set_profiling(aphpois* h, int on) {
  APH_CHECK_H(h);
  h->eng.set_profile(on != 0);
  return APHPOIS_OK;
}

int aphpois_kernel_times(aphpois* h, double* avg_ms, long long* count, int* nkernels) {
  APH_CHECK_H(h);
  const int nk = h->eng.prof_kernels();
  if (nkernels) *nkernels = nk;
  for (int j = 0; j < nk; ++j) {
    if (avg_ms) avg_ms[j] = h->eng.prof_avg_ms(j);
    if (count) count[j] = h->eng.prof_count(j);
  }
  return APHPOIS_OK;
}

int aphpois_launch_info(aphpois* h, int* stencil_blocks, int* stream_blocks, int* vx) {
  APH_CHECK_H(h);
  if (stencil_blocks) *stencil_blocks = h->eng.blocks_stencil();
  if (stream_blocks) *stream_blocks = h->eng.blocks_stream();
  if (vx) *vx = h->eng.vx();
  return APHPOIS_OK;
}

int aphpois_set_tile(aphpois* h, int lx, int bz, int vx, int sm) {
  APH_CHECK_H(h);
  if (lx < 0 || lx > 256 || (lx & (lx - 1)) || bz < 0 || vx < 0 || vx > 4 || sm < 0)
    return set_err(APHPOIS_EINVAL, "bad tile (lx a power of two <= 256, vx in 0,1,2,4)");
  aph::Tuning t;
  t.lx = lx;
  t.bz = bz;
  t.vx = vx;
  t.sm = sm;
  h->eng.set_tuning(t);
  return APHPOIS_OK;
}

int aphpois_assemble(aphpois* h, const double* rho, const double* ffv, const double* src,
                     double cell, double dt, int sign) {
  APH_CHECK_H(h);
  if (!rho) return set_err(APHPOIS_EINVAL, "rho is NULL");
  APH_RET(*h, h->eng.assemble(h->slot, rho, nullptr, ffv, src, cell, dt, sign, false));
}

int aphpois_assemble_device(aphpois* h, const double* rho, const double* ffv,
                            const double* src, double cell, double dt, int sign) {
  APH_CHECK_H(h);
  if (!rho) return set_err(APHPOIS_EINVAL, "rho is NULL");
  APH_RET(*h, h->eng.assemble(h->slot, rho, nullptr, ffv, src, cell, dt, sign, true));
}

int aphpois_assemble_faces(aphpois* h, const double* fk, const double* ffv, const double* src,
                           double cell, int sign) {
  APH_CHECK_H(h);
  if (!fk) return set_err(APHPOIS_EINVAL, "fk is NULL");
  APH_RET(*h, h->eng.assemble(h->slot, nullptr, fk, ffv, src, cell, 1.0, sign, false));
}

int aphpois_flux_correct(aphpois* h, const double* p, const double* rho, const double* ffv,
                         double* ffv_out, double cell, double dt) {
  APH_CHECK_H(h);
  if (!rho) return set_err(APHPOIS_EINVAL, "rho is NULL");
  APH_RET(*h, h->eng.flux_correct(h->slot, p, rho, nullptr, ffv, ffv_out, cell, dt, false));
}

int aphpois_flux_correct_device(aphpois* h, const double* p, const double* rho,
                                const double* ffv, double* ffv_out, double cell, double dt) {
  APH_CHECK_H(h);
  if (!rho) return set_err(APHPOIS_EINVAL, "rho is NULL");
  APH_RET(*h, h->eng.flux_correct(h->slot, p, rho, nullptr, ffv, ffv_out, cell, dt, true));
}

int aphpois_flux_correct_faces(aphpois* h, const double* p, const double* fk, const double* ffv,
                               double* ffv_out, double cell) {
  APH_CHECK_H(h);
  if (!fk) return set_err(APHPOIS_EINVAL, "fk is NULL");
  APH_RET(*h, h->eng.flux_correct(h->slot, p, nullptr, fk, ffv, ffv_out, cell, 1.0, false));
}

int aphpois_precond_apply(aphpois* h, const double* r, double* z) {
  APH_CHECK_H(h);
  if (!r || !z) return set_err(APHPOIS_EINVAL, "r or z is NULL");
  APH_RET(*h, h->eng.precond_apply(h->slot, r, z));
}

int aphpois_copy_bandwidth(int device, long long bytes, int reps, double* gbps) {
  if (bytes < 16 || reps < 1 || !gbps) return set_err(APHPOIS_EINVAL, "bad arguments");
  bytes &= ~15LL;
  if (hipSetDevice(device) != hipSuccess) return set_err(APHPOIS_EDEVICE, "no such device");
  void *a = nullptr, *b = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc(&a, bytes);
  if (e == hipSuccess) e = hipMalloc(&b, bytes);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipMemsetAsync(a, 0, bytes, st);
  if (e == hipSuccess) e = aph::launch_copy(a, b, bytes / 16, st);  // warm-up
  if (e == hipSuccess) e = hipEventRecord(e0, st);
  for (int r = 0; r < reps && e == hipSuccess; ++r) e = aph::launch_copy(a, b, bytes / 16, st);
  if (e == hipSuccess) e = hipEventRecord(e1, st);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e == hipSuccess) *gbps = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;
  if (e1) (void)hipEventDestroy(e1);
  if (e0) (void)hipEventDestroy(e0);
  if (st) (void)hipStreamDestroy(st);
  if (b) (void)hipFree(b);
  if (a) (void)hipFree(a);
  if (e != hipSuccess) return set_err(APHPOIS_EDEVICE, hipGetErrorString(e));
  return APHPOIS_OK;
}

int aphpois_destroy(aphpois* h) {
  delete h;
  return APHPOIS_OK;
}

// ---------------------------------------------------------------- batch

int aphpois_batch_create(const aphpois_batch_desc* d, aphpois_batch** out) {
  if (!d || !out) return set_err(APHPOIS_EINVAL, "NULL argument");
  *out = nullptr;
  aphpois_batch* b = new (std::nothrow) aphpois_batch();
  if (!b) return set_err(APHPOIS_ENOMEM, "host allocation failed");
  b->desc = *d;
  const int rc = b->eng.init(d->device, APHPOIS_PCG, d->precision, d->resmode, d->tol, 0,
                             d->maxiter, APHPOIS_NULL_AUTO, d->delta, d->check_every,
                             d->mem_budget, d->max_systems, d->precond, d->mg_w);
  if (rc != APHPOIS_OK) {
    set_err(rc, b->eng.err());
    delete b;
    return rc;
  }
  *out = b;
  return APHPOIS_OK;
}

int aphpois_batch_add(aphpois_batch* b, int nx, int ny, int nz, double h3, const double* expr,
                      const double* x0, int* slot) {
  APH_CHECK_H(b);
  if (!slot) return set_err(APHPOIS_EINVAL, "slot is NULL");
  int s = -1;
  int rc = b->eng.add_system(nx, ny, nz, h3, &s);
  if (rc == APHPOIS_OK) rc = b->eng.set_system(s, expr, false, 0, nx, (long long)nx * ny);
  if (rc == APHPOIS_OK) rc = b->eng.prepare(s, x0, false, 0, nx, (long long)nx * ny, false);
  if (rc != APHPOIS_OK) {
    set_err(rc, b->eng.err());
    if (s >= 0) b->eng.remove(s);
    return rc;
  }
  *slot = s;
  return APHPOIS_OK;
}

int aphpois_batch_set_rhs(aphpois_batch* b, int slot, const double* e7, const double* x0) {
  APH_CHECK_H(b);
  aph::System* s = b->eng.sys(slot);
  if (!s) return set_err(APHPOIS_EINVAL, "no such slot");
  int rc = APHPOIS_OK;
  if (e7) rc = b->eng.set_rhs(slot, e7, false, 0, s->nx, (long long)s->nx * s->ny);
  if (rc == APHPOIS_OK)
    rc = b->eng.prepare(slot, x0, false, 0, s->nx, (long long)s->nx * s->ny, x0 == nullptr);
  if (rc != APHPOIS_OK) return set_err(rc, b->eng.err());
  return APHPOIS_OK;
}

int aphpois_batch_solve(aphpois_batch* b) {
  APH_CHECK_H(b);
  APH_RET(*b, b->eng.run());
}

int aphpois_batch_get(aphpois_batch* b, int slot, double* x, aphpois_info* info) {
  APH_CHECK_H(b);
  aph::System* s = b->eng.sys(slot);
  if (!s) return set_err(APHPOIS_EINVAL, "no such slot");
  APH_RET(*b, b->eng.fetch(slot, x, false, 0, s->nx, (long long)s->nx * s->ny, info));
}

int aphpois_batch_remove(aphpois_batch* b, int slot) {
  APH_CHECK_H(b);
  APH_RET(*b, b->eng.remove(slot));
}

long long aphpois_batch_device_bytes(const aphpois_batch* b) { return b ? b->eng.bytes() : 0; }

int aphpois_batch_set_system_tile(aphpois_batch* b, int slot, int lx, int bz) {
  APH_CHECK_H(b);
  if (lx < 0 || lx > 256 || (lx & (lx - 1)) || bz < 0)
    return set_err(APHPOIS_EINVAL, "bad tile (lx a power of two <= 256)");
  APH_RET(*b, b->eng.set_system_tile(slot, lx, bz));
}

double aphpois_batch_last_seconds(const aphpois_batch* b) {
  return b ? b->eng.last_run_ms() * 1e-3 : 0.0;
}

int aphpois_batch_set_tile(aphpois_batch* b, int lx, int bz, int vx, int sm) {
  APH_CHECK_H(b);
  if (lx < 0 || lx > 256 || (lx & (lx - 1)) || bz < 0 || vx < 0 || vx > 4 || sm < 0)
    return set_err(APHPOIS_EINVAL, "bad tile (lx a power of two <= 256, vx in 0,1,2,4)");
  aph::Tuning t;
  t.lx = lx;
  t.bz = bz;
  t.vx = vx;
  t.sm = sm;
  b->eng.set_tuning(t);
  return APHPOIS_OK;
}

int aphpois_batch_destroy(aphpois_batch* b) {
  delete b;
  return APHPOIS_OK;
}

}  // extern "C"
