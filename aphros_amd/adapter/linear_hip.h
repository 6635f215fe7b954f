// linsolver_symm plugin "hip": the aphpois C ABI (include/aphpois.h) behind
// Aphros' own symmetric-solver interface.
//
//   linear::Solver<M>          src/linear/linear.h:15-57 (Conf, Info, Solve,
//                              SetConf, GetConf)
//   linear::ModuleLinear<M>    src/linear/linear.h:59-76 (Make, GetConf)
//   registration               src/linear/linear.cpp:16-23, src/util/module.h
//
// Compiled against the reference's headers; the product library it calls is
// libaphpois.so.  Solve() follows the staged protocol of the reference's
// device path (src/linear/conjugate_cl.ipp:36-138): every block copies its
// inner cells into the rank-wide field of the lead block, the lead block
// calls the C ABI once, every block copies the solution back and requests
// the direct_one halo exchange, and Info reaches every block through
// BcastFromLead.  The rank must own the whole domain (one GPU per domain);
// anything else is refused with fassert, like any other configuration error
// of the reference.
//
// Resident coefficients: while gathering, every block compares its cells'
// coefficients e0..e6 with those of the previous Solve bit for bit; when no
// block saw a change, the lead uploads only the constant term e7
// (aphpois_set_rhs: 8 instead of 64 bytes per cell) and the device keeps its
// converted operator and multigrid levels.  linsolver_hip_reuse = 0 turns
// this off.
//
// Walls: the C ABI treats every axis as periodic.  On a non-periodic axis the
// reference's pressure system links no cell across the domain boundary (the
// projection assembles wall and pressure boundary faces with a zero
// coefficient for the outer cell, src/solver/proj.ipp:343-351), so the
// periodic wrap only ever multiplies by zero and the solve is the same.  The
// gather checks that those coefficients are zero and the solve refuses the
// system otherwise.
#pragma once

#include <atomic>
#include <cstring>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>

#include "aphpois.h"
#include "linear/linear.h"
#include "util/logger.h"

namespace linear {

template <class M>
class SolverHip final : public Solver<M> {
 public:
  using Base = Solver<M>;
  using Conf = typename Base::Conf;
  using Info = typename Base::Info;
  using Scal = typename M::Scal;
  using Expr = typename M::Expr;
  using MIdx = typename M::MIdx;

  struct Extra {
    // residual of the stop rule: max-norm instead of L2 (CONJUGATE; the
    // reference's linsolver_<prefix>_maxnorm)
    int residual_max = 0;
    int method = APHPOIS_CONJUGATE;  // APHPOIS_CONJUGATE, APHPOIS_JACOBI or APHPOIS_PCG
    int precision = APHPOIS_FP64;    // APHPOIS_FP64 or APHPOIS_MIXED (PCG)
    int precond = APHPOIS_PRECOND_MG;  // PCG: APHPOIS_PRECOND_JACOBI or APHPOIS_PRECOND_MG
    // PCG: stop on the FP64 true relative residual ||b - A x|| / ||b|| <= tol
    // instead of the reference's volume-scaled recursive residual
    int true_relres = 0;
    int device = 0;
    long long mem_budget = 0;
    // keep the device operator when only e7 changed since the last Solve
    int reuse = 1;
  };

  SolverHip(const Conf& conf, const Extra& extra, const M& m)
      : Base(conf), extra_(extra) {
    if (m.IsLead()) {
      shared_obj_ = std::make_unique<Shared>();
      shared_ = shared_obj_.get();
    }
  }

  Info Solve(
      const FieldCell<Expr>& fc_system, const FieldCell<Scal>* fc_init,
      FieldCell<Scal>& fc_sol, M& m) override {
    auto sem = m.GetSem("hip-solve");
    struct {
      Info info{};
    } * ctx(sem);
    auto& t = *ctx;
    const M& ms = m.GetShared();
    if (sem()) {
      // the lead block sizes the rank-wide fields before any block writes
      // into them (the next stage)
      if (m.IsLead()) {
        Shared& s = *shared_;
        const size_t ns = ms.GetIndexCells().size();
        if (s.fcs.size() != ns) {
          s.fcs.Reinit(ms);
          s.resident = false;
        }
        if (s.fce7.size() != ns) s.fce7.Reinit(ms, Scal(0));
        if (s.fcu.size() != ns) s.fcu.Reinit(ms, Scal(0));
        s.changed = !s.resident || !extra_.reuse;
        s.wall_link = false;
      }
      m.BcastFromLead(&shared_);
    }
    if (sem("gather")) {
      Shared& s = *shared_;
      GatherSystem(fc_system, s, m);
      if (fc_init) {
        ToShared(*fc_init, s.fcu, m);
      }
    }
    if (sem("solve")) {
      if (m.IsLead()) {
        Shared& s = *shared_;
        const MIdx size = ms.GetInBlockCells().GetSize();
        fassert(
            size == ms.GetGlobalSize(),
            "linsolver hip: the rank must own the whole domain (one GPU per domain)");
        fassert(
            !s.wall_link.load(),
            "linsolver hip: a coefficient links a cell across a non-periodic domain "
            "boundary; the solver needs it to be zero there (walls and pressure "
            "boundaries of the projection are)");
        if (!s.h) {
          Create(s, ms);
          s.resident = false;
          s.changed = true;
        }
        Check(aphpois_set_conf(s.h, this->conf.tol, this->conf.miniter, this->conf.maxiter));
        const auto& ics = ms.GetIndexCells();
        const long long start = ics.GetIdx(ms.GetInBlockCells().GetBegin()).raw();
        const long long lead_y = ics.GetSize()[0];
        const long long lead_z = lead_y * ics.GetSize()[1];
        if (s.changed.load() || !s.resident) {
          // Expr is 8 contiguous Scal per cell slot: the AoS the C ABI reads
          s.resident = false;
          const double* expr = &s.fcs.data()[0][0];
          Check(aphpois_set_system(s.h, expr, start, lead_y, lead_z));
          s.resident = true;
          ++s.uploads;
        } else {
          Check(aphpois_set_rhs(s.h, s.fce7.data(), start, lead_y, lead_z));
          ++s.rhs_uploads;
        }
        aphpois_info info{};
        Check(aphpois_solve(
            s.h, fc_init ? s.fcu.data() : nullptr, s.fcu.data(), start, lead_y,
            lead_z, &info));
        t.info.residual = info.residual;
        t.info.iter = info.iter;
      }
      m.BcastFromLead(&t.info);
    }
    if (sem("result")) {
      FromShared(shared_->fcu, fc_sol, m);
      m.Comm(&fc_sol, M::CommStencil::direct_one);
      if (m.flags.linreport && m.IsRoot()) {
        std::cerr << std::scientific << "linear(hip) '" + fc_system.GetName() + "':"
                  << " res=" << t.info.residual << " iter=" << t.info.iter << std::endl;
      }
    }
    if (sem()) {
    }
    return t.info;
  }

 private:
  // rank-wide state, owned by the lead block
  struct Shared {
    aphpois* h = nullptr;
    FieldCell<Expr> fcs;   // the system of the last Solve
    FieldCell<Scal> fce7;  // its constant terms alone
    FieldCell<Scal> fcu;
    bool resident = false;  // the device holds the operator of fcs
    // set by any block (blocks gather concurrently) and read by the lead
    // in the next stage
    std::atomic<bool> changed{true};    // some e0..e6 differs from fcs
    std::atomic<bool> wall_link{false}; // nonzero coefficient across a wall
    long long uploads = 0, rhs_uploads = 0;
    ~Shared() {
      if (h) aphpois_destroy(h);
    }
  };

 public:
  // Full system uploads and constant-term-only uploads so far (lead block).
  long long GetUploads() const {
    return shared_obj_ ? shared_obj_->uploads : 0;
  }
  long long GetRhsUploads() const {
    return shared_obj_ ? shared_obj_->rhs_uploads : 0;
  }

 private:

  static void Check(int rc) {
    fassert(
        rc == APHPOIS_OK,
        "linsolver hip: " + std::string(aphpois_last_error()) + " (status " +
            std::to_string(rc) + ")");
  }

  void Create(Shared& s, const M& ms) {
    aphpois_desc d;
    aphpois_desc_default(&d);
    const MIdx size = ms.GetInBlockCells().GetSize();
    d.nx = size[0];
    d.ny = size[1];
    d.nz = size[2];
    d.h3 = ms.GetCellSize().prod();
    d.method = extra_.method;
    d.precision = extra_.precision;
    // the preconditioner is a PCG option
    d.precond = extra_.method == APHPOIS_PCG ? extra_.precond : APHPOIS_PRECOND_JACOBI;
    d.resmode = extra_.method == APHPOIS_PCG && extra_.true_relres
                    ? APHPOIS_RES_REL_TRUE
                    : (extra_.residual_max ? APHPOIS_RES_MAX : APHPOIS_RES_L2);
    d.tol = this->conf.tol;
    d.miniter = this->conf.miniter;
    d.maxiter = this->conf.maxiter;
    d.device = extra_.device;
    d.mem_budget = extra_.mem_budget;
    Check(aphpois_create(&d, &s.h));
  }

  // The block's system into the rank-wide fields (a disjoint region per
  // block): each cell's Expr, compared bit for bit with the one of the last
  // Solve, and its constant term; coefficients that link across a
  // non-periodic domain boundary must be zero.
  void GatherSystem(const FieldCell<Expr>& fl, Shared& s, const M& m) {
    const M& ms = m.GetShared();
    const auto& ics = ms.GetIndexCells();
    const MIdx gs = ms.GetGlobalSize();
    bool changed = false, wall = false;
    for (auto c : m.CellsM()) {
      const MIdx w(c);
      const size_t i = ics.GetIdx(w).raw();
      const Expr& e = fl[c];
      Expr& d = s.fcs.data()[i];
      if (!changed && std::memcmp(&d[0], &e[0], 7 * sizeof(Scal)) != 0) changed = true;
      d = e;
      s.fce7.data()[i] = e[7];
      for (size_t a = 0; a < M::dim; ++a) {
        if (ms.flags.is_periodic[a]) continue;
        if (w[a] == 0 && e[1 + 2 * a] != 0) wall = true;
        if (w[a] == gs[a] - 1 && e[2 + 2 * a] != 0) wall = true;
      }
    }
    if (changed) s.changed.store(true);
    if (wall) s.wall_link.store(true);
  }

  // Inner cells of this block into / out of the rank-wide field (each block
  // writes a disjoint region), as src/opencl/opencl.h:30-64 does.
  template <class T>
  static void ToShared(const FieldCell<T>& fl, FieldCell<T>& fs, const M& m) {
    const M& ms = m.GetShared();
    const auto& ics = ms.GetIndexCells();
    for (auto c : m.CellsM()) {
      fs[ics.GetIdx(MIdx(c))] = fl[c];
    }
  }
  template <class T>
  static void FromShared(const FieldCell<T>& fs, FieldCell<T>& fl, const M& m) {
    const M& ms = m.GetShared();
    if (fl.size() != m.GetIndexCells().size()) {
      fl.Reinit(m);
    }
    const auto& ics = ms.GetIndexCells();
    for (auto c : m.AllCellsM()) {
      const MIdx w(c);
      if (ms.GetInBlockCells().IsInside(w)) {
        fl[c] = fs[ics.GetIdx(w)];
      }
    }
  }

  Extra extra_;
  std::unique_ptr<Shared> shared_obj_;
  Shared* shared_ = nullptr;
};

// Module "hip" for linsolver_symm / linsolver_gen.  Configuration:
//   hypre_<prefix>_{tol,maxiter,miniter}   Conf, as every linear module
//   linsolver_<prefix>_maxnorm             max-norm residual (CONJUGATE)
//   linsolver_hip_method                   conjugate | jacobi | pcg (default conjugate)
//   linsolver_hip_precision                fp64 | mixed (default fp64)
//   linsolver_hip_precond                  jacobi | mg (default mg)
//   linsolver_hip_true_relres              1: PCG stops on ||b - A x|| / ||b|| <= tol
//   linsolver_hip_device                   HIP device ordinal (default 0)
//   linsolver_hip_reuse                    1 (default): upload only e7 when the
//                                          coefficients did not change
template <class M>
class ModuleLinearHip : public ModuleLinear<M> {
 public:
  ModuleLinearHip() : ModuleLinear<M>("hip") {}
  std::unique_ptr<Solver<M>> Make(const Vars& var, std::string prefix, const M& m) override {
    typename SolverHip<M>::Extra extra;
    extra.residual_max = var.Int("linsolver_" + prefix + "_maxnorm", 0);
    const std::string method = var.String("linsolver_hip_method", "conjugate");
    if (method == "conjugate") {
      extra.method = APHPOIS_CONJUGATE;
    } else if (method == "jacobi") {
      extra.method = APHPOIS_JACOBI;
    } else if (method == "pcg") {
      extra.method = APHPOIS_PCG;
    } else {
      fassert(false, "linsolver_hip_method: unknown method '" + method + "'");
    }
    const std::string prec = var.String("linsolver_hip_precision", "fp64");
    fassert(prec == "fp64" || prec == "mixed", "linsolver_hip_precision: fp64 or mixed");
    extra.precision = prec == "mixed" ? APHPOIS_MIXED : APHPOIS_FP64;
    const std::string pc = var.String("linsolver_hip_precond", "mg");
    fassert(pc == "mg" || pc == "jacobi", "linsolver_hip_precond: mg or jacobi");
    extra.precond = pc == "mg" ? APHPOIS_PRECOND_MG : APHPOIS_PRECOND_JACOBI;
    extra.true_relres = var.Int("linsolver_hip_true_relres", 0);
    extra.device = var.Int("linsolver_hip_device", 0);
    extra.reuse = var.Int("linsolver_hip_reuse", 1);
    return std::make_unique<SolverHip<M>>(this->GetConf(var, prefix), extra, m);
  }
};

}  // namespace linear
