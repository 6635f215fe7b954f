"""CPU checks of the product library that need no GPU: the C ABI surface
(exported symbols, struct layout), the synthetic-system generator against its
independent numpy twin, the assembly against the oracle, and the Python
mirror of the plugin interface."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from aphros_amd import _lib as L
from aphros_amd import linear, synth
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "aphpois.h")


def _declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(aphpois_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(L.LIB_PATH)
    names = _declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(L.SIGNATURES), set(names) ^ set(L.SIGNATURES)


def test_struct_layout_matches_c(tmp_path):
    src = tmp_path / "sz.c"
    src.write_text('#include "aphpois.h"\n#include <stdio.h>\n#include <stddef.h>\n'
                   "int main(void){printf(\"%zu %zu %zu %zu %zu\\n\", sizeof(aphpois_desc),"
                   " sizeof(aphpois_info), sizeof(aphpois_batch_desc),"
                   " offsetof(aphpois_desc, mem_budget), offsetof(aphpois_info, seconds));"
                   "return 0;}\n")
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, stdout=subprocess.PIPE, text=True).stdout
    got = [int(v) for v in out.split()]
    assert got == [C.sizeof(L.Desc), C.sizeof(L.Info), C.sizeof(L.BatchDesc),
                   L.Desc.mem_budget.offset, L.Info.seconds.offset]


def test_create_fails_cleanly_without_gpu():
    d = L.Desc()
    L.lib().aphpois_desc_default(C.byref(d))
    assert (d.method, d.miniter, d.maxiter, d.tol) == (L.CONJUGATE, 1, 100, 0.0)
    d.nx = d.ny = d.nz = 8
    h = C.c_void_p()
    rc = L.lib().aphpois_create(C.byref(d), C.byref(h))
    if rc == L.OK:
        L.lib().aphpois_destroy(h)
        pytest.skip("a GPU is present")
    assert rc in (L.EINVAL, L.EDEVICE)
    assert L.lib().aphpois_last_error().decode()


def test_invalid_descriptors_rejected():
    for field, value in (("method", 7), ("precision", 3), ("resmode", 9), ("maxiter", -1)):
        d = L.Desc()
        L.lib().aphpois_desc_default(C.byref(d))
        d.nx = d.ny = d.nz = 8
        setattr(d, field, value)
        h = C.c_void_p()
        assert L.lib().aphpois_create(C.byref(d), C.byref(h)) in (L.EINVAL, L.EDEVICE)


def test_noalloc_rejects_before_device_work():
    """The allocation-free solve refuses, on the host and before any device
    call, what its static work space or its kernels cannot take."""
    lib = L.lib()
    assert lib.aphpois_noalloc_max_cells() == 1 << 21
    info = L.Info()
    for shape, method, resmode in (((128, 128, 129), L.CONJUGATE, L.RES_L2),
                                   ((8, 8, 8), L.PCG, L.RES_L2),
                                   ((8, 8, 8), L.JACOBI, L.RES_REL_TRUE),
                                   ((0, 8, 8), L.CONJUGATE, L.RES_L2)):
        d = L.Desc()
        lib.aphpois_desc_default(C.byref(d))
        d.nz, d.ny, d.nx = shape
        d.method, d.resmode = method, resmode
        rc = lib.aphpois_solve_device_noalloc(C.byref(d), C.c_void_p(256), None,
                                              C.c_void_p(256), None, C.byref(info))
        assert rc == L.EINVAL, (shape, method, resmode)
    d = L.Desc()
    lib.aphpois_desc_default(C.byref(d))
    d.nx = d.ny = 128
    d.nz = 129
    lib.aphpois_solve_device_noalloc(C.byref(d), C.c_void_p(256), None, C.c_void_p(256), None,
                                     C.byref(info))
    assert b"2^21" in lib.aphpois_last_error()
    assert lib.aphpois_solve_device_noalloc(None, None, None, None, None, None) == L.EINVAL


@pytest.mark.parametrize("n,ns,rmin,rmax,ratio,seed", [
    (32, 8, 0.08, 0.15, 1000.0, 1), (48, 5, 0.05, 0.2, 10.0, 4), (20, 30, 0.02, 0.4, 3.0, 9),
    (17, 3, 0.3, 0.45, 1000.0, 123456789)])
def test_generator_matches_numpy_twin(n, ns, rmin, rmax, ratio, seed):
    rho = synth.density(n, ns, rmin, rmax, ratio, seed).reshape(n, n, n)
    assert np.array_equal(rho, O.twin_density(n, ns, rmin, rmax, ratio, seed))
    f = synth.rhs(n, seed).reshape(n, n, n)
    assert np.array_equal(f, O.twin_rhs(n, seed))
    assert abs(f.mean()) < 1e-15
    e = synth.assemble(n, rho, f).reshape(n, n, n, 8)
    assert np.array_equal(e, O.assemble(rho, f))


def test_assembled_system_invariants():
    e, rho, f = synth.config("c1_64")
    E = e.reshape(64, 64, 64, 8)
    assert (E[..., 0] > 0).all()
    assert np.array_equal(E[..., 1], np.roll(E[..., 2], 1, axis=2))
    assert np.array_equal(E[..., 3], np.roll(E[..., 4], 1, axis=1))
    assert np.array_equal(E[..., 5], np.roll(E[..., 6], 1, axis=0))
    rs = E[..., :7].sum(axis=-1)
    assert np.abs(rs).max() <= 1e-15 * E[..., 0].max()
    assert abs(E[..., 7].sum()) < 1e-12 * np.abs(E[..., 7]).sum()
    # A * 1 = 0 (periodic nullspace)
    assert np.abs(O.apply(E, np.ones((64, 64, 64)))).max() <= 1e-15 * E[..., 0].max()


def test_plugin_registry_and_conf():
    """The Python module "hip" reads the C++ module's keys with its defaults
    (aphros_amd/adapter/linear_hip.h ModuleLinearHip::Make)."""
    names = linear.ModuleLinear.GetInstances()
    assert "hip" in names
    var = linear.Vars(Int={"hypre_symm_maxiter": 50, "linsolver_symm_maxnorm": 1},
                      Double={"hypre_symm_tol": 1e-6},
                      String={"linsolver_symm": "hip"})
    conf = linear.ModuleLinear.GetConf(var, "symm")
    assert (conf.tol, conf.maxiter, conf.miniter) == (1e-6, 50, 0)
    m = linear.Mesh((8, 8, 8), hl=2)
    s = linear.MakeLinearSolver(var, "symm", m)
    assert (s.method, s.resmode, s.precision, s.reuse) == (L.CONJUGATE, L.RES_MAX, L.FP64, True)
    var.String["linsolver_hip_method"] = "pcg"
    var.String["linsolver_hip_precision"] = "mixed"
    var.Int["linsolver_hip_true_relres"] = 1
    var.Int["linsolver_hip_reuse"] = 0
    s = linear.MakeLinearSolver(var, "symm", m)
    assert (s.method, s.resmode, s.precision, s.precond, s.reuse) == \
        (L.PCG, L.RES_REL_TRUE, L.MIXED, L.PRECOND_MG, False)
    var.String["linsolver_hip_precond"] = "jacobi"
    assert linear.MakeLinearSolver(var, "symm", m).precond == L.PRECOND_JACOBI
    var.String["linsolver_symm"] = "nope"
    with pytest.raises(RuntimeError, match="Solver not found"):
        linear.MakeLinearSolver(var, "symm", m)
    var.String["linsolver_symm"] = "hip"
    var.String["linsolver_hip_precision"] = "fp16"
    with pytest.raises(RuntimeError):
        linear.MakeLinearSolver(var, "symm", m)
    var.String["linsolver_hip_precision"] = "fp64"
    var.String["linsolver_hip_method"] = "sor"
    with pytest.raises(RuntimeError, match="unknown method"):
        linear.MakeLinearSolver(var, "symm", m)


def test_python_and_cpp_modules_read_the_same_keys():
    """Every linsolver_hip_* key the C++ module reads is read by the Python
    module, and no other."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    cpp = open(os.path.join(root, "aphros_amd", "adapter", "linear_hip.h")).read()
    py = open(os.path.join(root, "aphros_amd", "linear.py")).read()
    keys = lambda t: set(re.findall(r'"(linsolver_hip_\w+)"', t))
    assert keys(cpp) == keys(py) and len(keys(cpp)) >= 6


def test_mesh_strides_and_halo_fill():
    m = linear.Mesh((5, 4, 3), hl=2)
    assert m.box() == (10, 9, 8)
    start, ly, lz = m.strides()
    assert (ly, lz) == (10, 90) and start == 2 * (1 + 10 + 90)
    inner = np.arange(60, dtype=float).reshape(3, 4, 5)
    fc = np.zeros(m.nslots())
    a = fc.reshape(8, 9, 10)
    a[2:5, 2:6, 2:7] = inner
    m.fill_halos(fc)
    padded = np.pad(inner, 2, mode="wrap")
    assert np.array_equal(a[:7, :8, :9], padded)
