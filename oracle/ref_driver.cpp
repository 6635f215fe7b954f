// Harness that runs the reference's own linear::SolverConjugate / SolverJacobi
// templates (src/linear/linear.ipp:24-129,158-217) on one block that covers the
// whole periodic grid.  Built only by oracle/Makefile.ref from the reference's
// sources; its outputs are test fixtures, never part of the product.
//
// The stage loop mirrors DistrMesh::Run (src/distr/distr.ipp:508-583) for a
// single block: call Solve once per stage, then serve the Comm requests with a
// periodic halo fill (as Local::ReadBuffer, src/distr/local.ipp:320-346) and
// drop the Reduce requests (one block: the per-block value is the total,
// src/distr/local.ipp:153-172).
//
// usage: ref_driver <conjugate|jacobi> nx ny nz tol miniter maxiter maxnorm
//                   expr.bin x0.bin|- x_out.bin
// expr.bin: nx*ny*nz*8 doubles, x-fastest cells, Expr components e0..e7.
// Prints "iter=<int> res=<%.17e> sec=<solve wall seconds>" on stdout (sec:
// the stage loop only, without file input and output).
//
// Built with -DAPH_HIP_MODULE (ref_hip_driver, linked with libaphpois and the
// "hip" module of aphros_amd/adapter): the solver is created through the
// reference's module registry, ModuleLinear<M>::GetInstance(name)->Make(var,
// "symm", m) (src/util/linear.ipp:9-28), so name may also be "hip"; extra
// arguments key=value set Vars (e.g. linsolver_hip_method=pcg).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#ifdef APH_HIP_MODULE
// the solver templates are instantiated by the reference's linear.cpp
#include "linear/linear.h"
#include "parse/vars.h"
#else
#include "linear/linear.ipp"
#endif

using M = MeshCartesian<double, 3>;
using Scal = double;
using MIdx = typename M::MIdx;
using Vect = typename M::Vect;
using Expr = typename M::Expr;

static std::vector<double> ReadAll(const std::string& path, size_t count) {
  std::vector<double> v(count);
  std::ifstream f(path, std::ios::binary);
  if (!f.read(reinterpret_cast<char*>(v.data()), count * sizeof(double))) {
    std::fprintf(stderr, "cannot read %zu doubles from %s\n", count, path.c_str());
    std::exit(2);
  }
  return v;
}

static void FillHalos(FieldCell<Scal>& fc, const M& m, MIdx gs) {
  auto& ic = m.GetIndexCells();
  for (auto c : m.AllCells()) {
    MIdx w = ic.GetMIdx(c);
    MIdx wr = w;
    for (int d = 0; d < 3; ++d) {
      wr[d] = ((w[d] % gs[d]) + gs[d]) % gs[d];
    }
    if (!(wr == w)) {
      fc[c] = fc[ic.GetIdx(wr)];
    }
  }
}

int main(int argc, char** argv) {
#ifdef APH_HIP_MODULE
  if (argc < 12) {
#else
  if (argc != 12) {
#endif
    std::fprintf(stderr, "bad args\n");
    return 2;
  }
  const std::string name = argv[1];
  const MIdx gs(std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]));
  typename linear::Solver<M>::Conf conf;
  conf.tol = std::atof(argv[5]);
  conf.miniter = std::atoi(argv[6]);
  conf.maxiter = std::atoi(argv[7]);
  const int maxnorm = std::atoi(argv[8]);
  const size_t ncells = size_t(gs[0]) * gs[1] * gs[2];

  const Rect<Vect> domain(Vect(0), Vect(gs) / Vect(gs.max()));
  M m(MIdx(0), gs, domain, 2, true, true, gs, 0);

  auto expr = ReadAll(argv[9], ncells * 8);
  FieldCell<Expr> fcs(m, Expr(0));
  FieldCell<Scal> sol(m, 0);
  const bool has_x0 = std::string(argv[10]) != "-";
  std::vector<double> x0;
  if (has_x0) {
    x0 = ReadAll(argv[10], ncells);
  }
  auto& ic = m.GetIndexCells();
  for (auto c : m.Cells()) {
    const MIdx w = ic.GetMIdx(c);
    const size_t i = size_t(w[0]) + size_t(gs[0]) * (w[1] + size_t(gs[1]) * w[2]);
    for (int k = 0; k < 8; ++k) {
      fcs[c][k] = expr[i * 8 + k];
    }
    if (has_x0) {
      sol[c] = x0[i];
    }
  }
  FillHalos(sol, m, gs);

  std::unique_ptr<linear::Solver<M>> solver;
#ifdef APH_HIP_MODULE
  m.SetShared(&m);
  for (int d = 0; d < 3; ++d) m.flags.is_periodic[d] = true;
  Vars var;
  var.Double.Set("hypre_symm_tol", conf.tol);
  var.Int.Set("hypre_symm_miniter", conf.miniter);
  var.Int.Set("hypre_symm_maxiter", conf.maxiter);
  var.Int.Set("linsolver_symm_maxnorm", maxnorm);
  for (int a = 12; a < argc; ++a) {
    const std::string kv = argv[a];
    const size_t eq = kv.find('=');
    if (eq == std::string::npos) {
      std::fprintf(stderr, "bad key=value %s\n", argv[a]);
      return 2;
    }
    const std::string key = kv.substr(0, eq), val = kv.substr(eq + 1);
    if (key.find("method") != std::string::npos || key.find("precision") != std::string::npos ||
        key.find("precond") != std::string::npos) {
      var.String.Set(key, val);
    } else {
      var.Int.Set(key, std::atoi(val.c_str()));
    }
  }
  auto* mod = linear::ModuleLinear<M>::GetInstance(name);
  if (!mod) {
    std::fprintf(stderr, "no linear module %s\n", name.c_str());
    return 2;
  }
  solver = mod->Make(var, "symm", m);
#else
  if (name == "conjugate") {
    typename linear::SolverConjugate<M>::Extra extra;
    extra.residual_max = maxnorm;
    solver.reset(new linear::SolverConjugate<M>(conf, extra, m));
  } else if (name == "jacobi") {
    typename linear::SolverJacobi<M>::Extra extra;
    solver.reset(new linear::SolverJacobi<M>(conf, extra, m));
  } else {
    std::fprintf(stderr, "unknown solver %s\n", name.c_str());
    return 2;
  }
#endif

  typename linear::Solver<M>::Info info{};
  const auto t0 = std::chrono::steady_clock::now();
  do {
    info = solver->Solve(fcs, has_x0 ? &sol : nullptr, sol, m);
    for (auto& req : m.GetComm()) {
      auto* r = dynamic_cast<typename M::CommRequestScal*>(req.get());
      if (!r) {
        std::fprintf(stderr, "unexpected comm request\n");
        return 2;
      }
      FillHalos(*r->field, m, gs);
    }
    m.ClearComm();
    m.ClearReduce();
    m.ClearBcastFromLead();  // one block: the lead's values are in place
  } while (m.Pending());
  const double sec =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

  std::vector<double> out(ncells);
  for (auto c : m.Cells()) {
    const MIdx w = ic.GetMIdx(c);
    out[size_t(w[0]) + size_t(gs[0]) * (w[1] + size_t(gs[1]) * w[2])] = sol[c];
  }
  std::ofstream f(argv[11], std::ios::binary);
  f.write(reinterpret_cast<const char*>(out.data()), ncells * sizeof(double));
  std::printf("iter=%d res=%.17e sec=%.6f\n", info.iter, info.residual, sec);
  return 0;
}
