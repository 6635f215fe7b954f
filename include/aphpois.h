/*
 * aphpois — MI355X-native pressure-Poisson solver, C ABI.
 *
 * Drop-in for the symmetric linear-solver plugin of Aphros (`linsolver_symm`,
 * src/linear/linear.h:15-57).  The system is the per-cell 7-point Expr
 *     e0*x[c] + e1*x[xm] + e2*x[xp] + e3*x[ym] + e4*x[yp] + e5*x[zm] + e6*x[zp]
 *       + e7 = 0
 * (src/linear/linear.h:37-39, src/geom/mesh.h:482-485), neighbour order
 * q = 2*d + side (src/geom/mesh.ipp:100-107).  Plain pointers and sizes only.
 *
 * Every function returns APHPOIS_OK (0) or an error code; the message of the
 * last error on the calling thread is returned by aphpois_last_error().
 * Nothing aborts: a failing call leaves other handles untouched.
 */
#ifndef APHPOIS_H
#define APHPOIS_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
enum {
  APHPOIS_OK = 0,
  APHPOIS_EINVAL = 1,   /* bad argument or unsupported system */
  APHPOIS_ENOMEM = 2,   /* device allocation failed */
  APHPOIS_EDEVICE = 3,  /* HIP runtime error */
  APHPOIS_ESTATE = 4,   /* call out of order (e.g. solve before set_system) */
  APHPOIS_EBUDGET = 5,  /* memory budget would be exceeded */
  APHPOIS_ENUMERIC = 6, /* NaN/Inf in the system or breakdown in the solve */
  APHPOIS_ECANCEL = 7   /* request cancelled */
};

/* ---- method: which solver runs ---- */
enum {
  /* SolverConjugate restated on the GPU: unpreconditioned CG, FP64, the
   * reference stop rule (src/linear/linear.ipp:24-129). */
  APHPOIS_CONJUGATE = 0,
  /* SolverJacobi restated on the GPU (src/linear/linear.ipp:158-217). */
  APHPOIS_JACOBI = 1,
  /* Preconditioned CG (Jacobi scaling or multigrid) with reliable updates in
   * FP64 (new; symmetric systems with a same-sign diagonal).  The stop test
   * runs at every FP64 true-residual (reliable) update: it stops once
   * iter >= miniter and (the residual meets tol or iter > maxiter), the
   * reference's rule, so at most maxiter + 1 iterations; unlike the
   * reference it may stop at iteration 0 when the initial guess already
   * meets tol and miniter is 0. */
  APHPOIS_PCG = 2
};

/* ---- precision of the PCG inner iteration ---- */
enum {
  APHPOIS_FP64 = 0,
  /* FP32 face conductances, scale and Krylov vectors; FP64 true residual,
   * FP64 reductions, FP64 solution and mean removal. */
  APHPOIS_MIXED = 1
};

/* ---- residual / stop rule ---- */
enum {
  /* sqrt(sum r^2 / h3) < tol  (src/linear/linear.ipp:103-111) */
  APHPOIS_RES_L2 = 0,
  /* max|r| / h3 < tol  (linsolver_symm_maxnorm = 1) */
  APHPOIS_RES_MAX = 1,
  /* ||b - A x||_2 / ||b||_2 <= tol, b = -e7, FP64 (PCG only) */
  APHPOIS_RES_REL_TRUE = 2
};

/* ---- preconditioner of PCG ---- */
enum {
  /* diagonal (Jacobi) scaling; one streaming SpMV per iteration */
  APHPOIS_PRECOND_JACOBI = 0,
  /* one symmetric multigrid V-cycle per iteration (2x2x2 aggregation with
   * the sum of the covered face conductances, damped-Jacobi smoothing);
   * every grid edge must have enough factors of 2 to coarsen to <= 8192
   * cells (FP64: 4096) */
  APHPOIS_PRECOND_MG = 1
};

/* ---- nullspace handling for PCG ---- */
enum {
  APHPOIS_NULL_AUTO = 0, /* remove the mean when the system is singular */
  APHPOIS_NULL_OFF = 1,
  APHPOIS_NULL_ON = 2
};

/* W levels of the multigrid cycle on the largest grids: at 512^3 (9 levels)
 * the mixed-precision solve to 1e-8 takes 117 iterations with a V-cycle and
 * 35 with three W levels, at a third more work per iteration. */
#define APHPOIS_MG_W_DEFAULT 3
/* mg_w chosen per system from its cell count n (the default): a W level's
 * extra coarse cycles cost launch latency, the iterations it saves cost
 * bandwidth, so 0 for n < 2^22.5, 1 below 2^24.5, 2 below 2^25.5, else 3 */
#define APHPOIS_MG_W_AUTO (-1)

typedef struct aphpois_desc {
  int nx, ny, nz;   /* inner cells; the grid is periodic in x, y and z */
  double h3;        /* GetCellSize().prod(), used by the reference norms */
  int method;       /* APHPOIS_CONJUGATE | APHPOIS_JACOBI | APHPOIS_PCG */
  int precision;    /* APHPOIS_FP64 | APHPOIS_MIXED (PCG only) */
  int resmode;      /* APHPOIS_RES_* */
  double tol;       /* Conf::tol */
  int miniter;      /* Conf::miniter */
  int maxiter;      /* Conf::maxiter */
  int nullspace;    /* APHPOIS_NULL_* (PCG only) */
  double delta;     /* reliable-update trigger (PCG), 0 = default 0.05 */
  int check_every;  /* iterations between host polls, 0 = auto */
  int device;       /* HIP device ordinal */
  long long mem_budget; /* bytes of device memory this handle may use, 0 = no limit */
  int precond;      /* PCG preconditioner: APHPOIS_PRECOND_* */
  int mg_w;         /* multigrid: the top mg_w levels apply their coarse
                       correction twice (0 = V-cycle); default
                       APHPOIS_MG_W_AUTO */
} aphpois_desc;

typedef struct aphpois_info {
  int iter;           /* Info::iter */
  double residual;    /* Info::residual in the handle's resmode */
  double res_l2;      /* sqrt(sum r^2 / h3) of the final residual */
  double res_max;     /* max|r| / h3 of the final residual */
  double true_relres; /* ||b - A x|| / ||b|| recomputed in FP64 at the end */
  double mean;        /* mean removed from x (PCG with nullspace removal) */
  int converged;      /* 1 if the stop rule was met before maxiter */
  int refinements;    /* PCG: FP64 true-residual (reliable) updates */
  int status;         /* APHPOIS_OK or APHPOIS_ENUMERIC on breakdown */
  double seconds;     /* wall time of the solve on the device stream */
  double max_diff;    /* max |x - xref| against the reference solution set
                         by aphpois_set_solution_ref, the mean of x - xref
                         removed when the system is singular (the max_diff of
                         the reference's linear test,
                         src/test/linear/main.cpp:107-128); -1 if none */
} aphpois_info;

typedef struct aphpois aphpois;

/* Fills *d with defaults: method CONJUGATE, FP64, RES_L2, tol 0, miniter 1,
 * maxiter 100 (linear::Solver::Conf defaults, src/linear/linear.h:21-25). */
void aphpois_desc_default(aphpois_desc* d);

/* Creates a solver handle (replaces ModuleLinear::Make / the Solver
 * constructor, src/linear/linear.ipp:243-252). */
int aphpois_create(const aphpois_desc* desc, aphpois** out);

/* Uploads the system.  expr points to host memory holding 8 doubles per slot
 * (Expr, AoS); cell (ix,iy,iz) of the inner grid is at slot
 *   start + iz*lead_z + iy*lead_y + ix
 * exactly as d_system is addressed by the reference's OpenCL kernel
 * (src/linear/conjugate_cl.cl:14-18, src/opencl/opencl.ipp:310-312).
 * For a dense x-fastest array pass start=0, lead_y=nx, lead_z=nx*ny. */
int aphpois_set_system(aphpois* h, const double* expr, long long start,
                       long long lead_y, long long lead_z);

/* Same, with expr in device memory (dense x-fastest, 8 doubles per cell). */
int aphpois_set_system_device(aphpois* h, const double* expr_dev);

/* Replaces only the constant term e7 (host, same strides as set_system,
 * one double per slot).  Coefficients stay resident. */
int aphpois_set_rhs(aphpois* h, const double* e7, long long start,
                    long long lead_y, long long lead_z);

/* Solves (replaces Solver::Solve, src/linear/linear.h:45-47).  x0 may be NULL
 * (zero guess) or equal x (warm start); both are host scalar fields with the
 * strides given; only inner cells are read and written. */
int aphpois_solve(aphpois* h, const double* x0, double* x, long long start,
                  long long lead_y, long long lead_z, aphpois_info* info);

/* Same with dense device arrays (x0_dev may be NULL or equal x_dev). */
int aphpois_solve_device(aphpois* h, const double* x0_dev, double* x_dev,
                         aphpois_info* info);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): aphpois_create + aphpois_set_system_device +
 * aphpois_solve_device in one call that allocates nothing.  The same kernels,
 * tilings and reductions, so x and info equal that path's bit for bit.
 * desc: method CONJUGATE or JACOBI, resmode RES_L2 or RES_MAX, nx, ny, nz,
 * h3, tol, miniter, maxiter, device (other fields ignored).  expr_dev: dense
 * Expr in device memory (8 doubles per cell, x fastest), read in place;
 * x0_dev: NULL (zero guess), x_dev, or another device array; x_dev: the
 * solution.  stream: a hipStream_t of that device, NULL for its null stream;
 * the call returns after the solve has finished on it.  Work space (tables,
 * partials, work vectors) is one static device array per device, held until
 * the process ends: at most aphpois_noalloc_max_cells() = 2^21 cells
 * (EINVAL beyond).  Calls that overlap on one device share that array and
 * race; info.mean is 0 and info.max_diff -1. */
int aphpois_solve_device_noalloc(const aphpois_desc* desc, const double* expr_dev,
                                 const double* x0_dev, double* x_dev, void* stream,
                                 aphpois_info* info);
long long aphpois_noalloc_max_cells(void);

/* Sets the reference solution that every later solve of the handle is
 * compared with (aphpois_info.max_diff): an FP64 solution of the resident
 * system, e.g. of an FP64 solve or the exact solution of a test problem
 * (the u* of src/test/linear/main.cpp:45-60).  Host scalar field with the
 * strides of aphpois_solve; xref NULL removes it.  Kept in device memory
 * (8 bytes per cell, counted in the budget). */
int aphpois_set_solution_ref(aphpois* h, const double* xref, long long start,
                             long long lead_y, long long lead_z);

/* Replaces Solver::SetConf (src/linear/linear.h:48-50). */
int aphpois_set_conf(aphpois* h, double tol, int miniter, int maxiter);

/* Bytes of device memory held by the handle. */
long long aphpois_device_bytes(const aphpois* h);

int aphpois_destroy(aphpois* h);

/* ---------------- the projection around the solve ----------------------- */

/* Assembles the projection's pressure system on the device and makes it the
 * resident system, as aphpois_set_system would with the Expr that
 * Proj::GetFlux / GetFluxSum build (src/solver/proj.ipp:343-383; harmonic
 * face density src/solver/approx_eb.h:351-363; AppendExpr
 * src/geom/mesh.h:575-579), bit for bit: per face f between cells cm < cp,
 * ExprFace(-1/h, 1/h, 0) * (-area / rho_f * dt) + (0, 0, ffv[f]), summed
 * over each cell's faces with the outward factor, minus src * volume.
 * rho: cell density (nx*ny*nz); ffv: face volume fluxes, three arrays of
 * nx*ny*nz (x, y, z faces; entry c of direction d is the face between c - e_d
 * and c); src: volume source per cell, or NULL; cell: the cell size h (cubic
 * cells); sign +1 for the projection's convention (e0 > 0), -1 for the
 * negated one.  Dense host arrays, x fastest; _device takes device
 * pointers. */
int aphpois_assemble(aphpois* h, const double* rho, const double* ffv, const double* src,
                     double cell, double dt, int sign);
int aphpois_assemble_device(aphpois* h, const double* rho, const double* ffv,
                            const double* src, double cell, double dt, int sign);
/* The same with the face denominator given per face instead of the harmonic
 * density: ffe = ExprFace(-1/h, 1/h, 0) * (-area / fk[f]) + (0, 0, ffv[f]),
 * the pressure-correction system of the SIMPLE solver
 * (SIMPLE::GetFlux / GetFluxSum, src/solver/simple.ipp:214-240,359-372; fk
 * is its face momentum coefficient ffk).  fk: three face arrays like ffv. */
int aphpois_assemble_faces(aphpois* h, const double* fk, const double* ffv, const double* src,
                           double cell, int sign);
/* Corrected face fluxes for the pressure p (Proj::Project stage "fluxes",
 * src/solver/proj.ipp:399-403; UEB::Eval, src/solver/approx_eb.h:338-344):
 * ffv_out[f] = p[cm] * ffe[0] + p[cp] * ffe[1] + ffe[2], with ffe of the
 * assembly above.  Same layouts; ffv_out may not alias ffv. */
int aphpois_flux_correct(aphpois* h, const double* p, const double* rho, const double* ffv,
                         double* ffv_out, double cell, double dt);
int aphpois_flux_correct_device(aphpois* h, const double* p, const double* rho,
                                const double* ffv, double* ffv_out, double cell, double dt);
/* The SIMPLE form (src/solver/simple.ipp:384-387): face denominators fk. */
int aphpois_flux_correct_faces(aphpois* h, const double* p, const double* fk, const double* ffv,
                               double* ffv_out, double cell);

/* ---------------- system files ------------------------------------------ */

/* The linear system as one array of shape (nz, ny, nx, 8), float64, C order:
 * the dataset of the reference's --system_in / --system_out
 * (src/test/linear/main.cpp:181-189; HDF5 there), stored as a NumPy .npy
 * file (readable by numpy.load without pickles).  No device needed. */
int aphpois_system_file_info(const char* path, int* nx, int* ny, int* nz);
/* count must be nx*ny*nz*8. */
int aphpois_system_file_read(const char* path, double* expr, long long count);
int aphpois_system_file_write(const char* path, int nx, int ny, int nz, const double* expr);
/* Message of the last failing system-file call on this thread. */
const char* aphpois_system_file_error(void);

/* Message of the last failing call on this thread ("" if none). */
const char* aphpois_last_error(void);

/* ---------------- measurement hooks (no reference counterpart) ----------- */

/* Profiling mode: the iteration kernels are launched one by one on the
 * handle's stream with a HIP event before each (instead of replaying the
 * captured graph), so that per-kernel device time can be read back.  Off by
 * default. */
int aphpois_set_profiling(aphpois* h, int on);
/* Average device time (ms) and count of each iteration kernel over the
 * iterations that ran since profiling was switched on.  Kernel order (each
 * reducing kernel is followed by its finalize kernel): PCG {spmv, fin,
 * update, fin, ru_accum, fin, ru_resid, fin}, where the update's finalize
 * and the reliable-update kernels are timed together in the fourth entry
 * (the fifth to eighth stay empty), and, with the multigrid
 * preconditioner, {V-cycle (all its launches together), level-0 sweep (per
 * launch), level-0 restriction, coarser levels (per cycle), bottom levels,
 * level-0 pre-sweep, level-0 first post-sweep, level-0 second post-sweep};
 * CONJUGATE {spmv, fin, update, fin}; JACOBI {sweep, fin}.  avg_ms and count
 * hold 16 entries; *nkernels receives the number used. */
int aphpois_kernel_times(aphpois* h, double* avg_ms, long long* count, int* nkernels);
/* Launch geometry of the last solve: workgroups of the stencil and streaming
 * launches and the x vector width. */
int aphpois_launch_info(aphpois* h, int* stencil_blocks, int* stream_blocks, int* vx);
/* Overrides the launch shape (0 = automatic for each): lx threads per tile
 * row (power of two <= 256), bz planes per stencil workgroup, vx cells per
 * thread in x (1, 2 or 4; 4 only in MIXED; must divide nx), sm cells per
 * thread of the streaming kernels. */
int aphpois_set_tile(aphpois* h, int lx, int bz, int vx, int sm);
/* Practical HBM bandwidth of a plain streaming copy kernel on `device`:
 * copies `bytes` (two buffers of that size are allocated and freed) `reps`
 * times after one warm-up copy; *gbps = 2 * bytes * reps / time. */
int aphpois_copy_bandwidth(int device, long long bytes, int reps, double* gbps);
/* z = M r: one application of the multigrid preconditioner (method PCG,
 * precond MG) of the resident system, in the handle's precision; r and z are
 * dense host arrays of nx*ny*nz doubles, x fastest, in the sign convention of
 * the positive definite form s*A (s = sign of e0).  A diagnostic: it lets the
 * V-cycle be checked against an independent implementation. */
int aphpois_precond_apply(aphpois* h, const double* r, double* z);

/* ---------------- batched solves (many independent systems) -------------- */

typedef struct aphpois_batch aphpois_batch;

typedef struct aphpois_batch_desc {
  int precision;        /* APHPOIS_FP64 | APHPOIS_MIXED; method is PCG */
  int resmode;          /* APHPOIS_RES_REL_TRUE recommended */
  double tol;
  int maxiter;
  double delta;         /* 0 = default */
  int check_every;      /* 0 = auto */
  int device;
  long long mem_budget; /* bytes, 0 = no limit */
  int max_systems;      /* slot capacity, 0 = 1024 */
  int precond;          /* APHPOIS_PRECOND_* */
  int mg_w;             /* as in aphpois_desc (AUTO: by each system's size);
                           a batch shares its launches among its systems, so
                           its W levels cost little latency per system: 2 suit
                           batches of systems up to ~160^3 */
} aphpois_batch_desc;

int aphpois_batch_create(const aphpois_batch_desc* d, aphpois_batch** out);
/* Adds a periodic system (dense host Expr AoS, x-fastest).  x0 may be NULL.
 * Returns the slot id in *slot. */
int aphpois_batch_add(aphpois_batch* b, int nx, int ny, int nz, double h3,
                      const double* expr, const double* x0, int* slot);
/* Replaces the constant term of a resident system (dense host, one double per
 * cell) and, if x0 is not NULL, its initial guess; if x0 is NULL the previous
 * solution is the initial guess (warm start). */
int aphpois_batch_set_rhs(aphpois_batch* b, int slot, const double* e7,
                          const double* x0);
/* Solves every resident system that is not yet converged, all in the same
 * launches. */
int aphpois_batch_solve(aphpois_batch* b);
/* Copies a system's solution (dense host) and info. */
int aphpois_batch_get(aphpois_batch* b, int slot, double* x, aphpois_info* info);
int aphpois_batch_remove(aphpois_batch* b, int slot);
long long aphpois_batch_device_bytes(const aphpois_batch* b);
/* Overrides the launch shape for the batch (see aphpois_set_tile). */
int aphpois_batch_set_tile(aphpois_batch* b, int lx, int bz, int vx, int sm);
/* Stencil tile of one system of the batch (chosen by the autotuner). */
int aphpois_batch_set_system_tile(aphpois_batch* b, int slot, int lx, int bz);
/* Wall time (s) of the last aphpois_batch_solve on the device stream. */
double aphpois_batch_last_seconds(const aphpois_batch* b);
int aphpois_batch_destroy(aphpois_batch* b);

/* ------------- request server (continuous batching, one GPU) ------------- */
/* No counterpart in the reference: the reference solves one system per call
 * from inside the time loop.  The server keeps every admitted system's
 * coefficients and Krylov state resident in device memory, admits requests
 * as slots and memory free up, and advances all resident systems in the same
 * kernel launches. */

typedef struct aphpois_server aphpois_server;

typedef struct aphpois_server_desc {
  int precision;        /* APHPOIS_FP64 | APHPOIS_MIXED (method PCG) */
  double tol;           /* true relative residual per time step */
  int maxiter;          /* per time step */
  double delta;         /* reliable-update trigger, 0 = default */
  int check_every;      /* iterations per graph replay, 0 = auto */
  int device;
  long long mem_budget; /* admission budget in bytes of device memory, 0 = none */
  int max_active;       /* resident systems at most, 0 = 1024 */
  int precond;          /* APHPOIS_PRECOND_* */
  int mg_w;             /* as in aphpois_desc */
} aphpois_server_desc;

typedef struct aphpois_request {
  int nx, ny, nz;
  double h3;
  const double* expr;     /* host Expr AoS, dense; read at admission */
  int nsteps;             /* time steps solved in sequence, each warm-started
                             from the previous step's solution */
  const double* const* e7;/* optional: e7[k] = host constant term of step k
                             (k >= 1; NULL entries or NULL array: see perturb) */
  double perturb;         /* step k+1 without e7: b += perturb * max|b| * eta,
                             eta zero-mean uniform in [-1, 1] */
  unsigned long long seed;/* stream of eta */
  double* x_out;          /* host dense solution of the last step, or NULL */
} aphpois_request;

typedef struct aphpois_result {
  long long id;
  int status;             /* APHPOIS_OK, ECANCEL, EBUDGET, ENUMERIC, ... */
  int steps_done;
  long long iters_total;  /* inner iterations over all steps */
  int refinements_total;
  double last_relres;     /* true relative residual of the last step */
  double t_submit, t_admit, t_done;  /* seconds since server creation */
} aphpois_result;

typedef struct aphpois_server_stats {
  int queued, active;
  long long completed;
  long long device_bytes;
  double time;
} aphpois_server_stats;

int aphpois_server_create(const aphpois_server_desc* d, aphpois_server** out);
/* Thread-safe and non-blocking.  The request and the arrays it points to must
 * stay valid until the request completes. */
int aphpois_server_submit(aphpois_server* s, const aphpois_request* r, long long* id);
/* Waits for completion (timeout_s < 0: no limit) and releases the id. */
int aphpois_server_wait(aphpois_server* s, long long id, double timeout_s,
                        aphpois_result* out);
int aphpois_server_cancel(aphpois_server* s, long long id);
int aphpois_server_get_stats(aphpois_server* s, aphpois_server_stats* st);
/* Cancels queued requests, stops the worker and frees the device state. */
int aphpois_server_destroy(aphpois_server* s);
const char* aphpois_server_last_error(void);

/* ------------------------- synthetic systems (host) ---------------------- */

/* Density field of the benchmark systems: unit periodic cube, n^3 cells,
 * `nspheres` spheres with centres uniform in [0,1)^3 and radii uniform in
 * [rmin, rmax] drawn from splitmix64 streams of `seed`; rho = 1 inside,
 * `ratio` outside (periodic distance). */
int aphpois_synth_density(int n, int nspheres, double rmin, double rmax,
                          double ratio, unsigned long long seed, double* rho);
/* RHS f uniform in [-1,1] from the splitmix64 stream of `seed`, mean removed. */
int aphpois_synth_rhs(int n, unsigned long long seed, double* f);
/* Pressure Expr (AoS, dense) from cell density and volume source with the
 * reference's harmonic face density and projection sign
 * (src/solver/proj.ipp:343-383, src/solver/approx_eb.h:351-363), h = 1/n,
 * dt = 1, e7 = -f*h^3. */
int aphpois_synth_assemble(int n, const double* rho, const double* f,
                           double* expr);

#ifdef __cplusplus
}
#endif

#endif /* APHPOIS_H */
