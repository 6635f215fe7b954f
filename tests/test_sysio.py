"""System files (aphpois_system_file_*): the (nz, ny, nx, 8) float64 dataset
of the reference's --system_in / --system_out, as .npy.  Checked against
numpy's own .npy reader and writer (CPU only)."""
import numpy as np
import pytest

from aphros_amd import _lib as L
from aphros_amd import sysio
from oracle import oracle as O


def test_roundtrip_with_numpy(tmp_path):
    E = O.tlinear_system(8)[0]
    E = np.ascontiguousarray(E[:, :6, :5])  # nz != ny != nx
    p = str(tmp_path / "sys.npy")
    sysio.save_system(p, E)
    assert np.array_equal(np.load(p, allow_pickle=False), E)
    q = str(tmp_path / "np.npy")
    np.save(q, E * 3.0)
    assert np.array_equal(sysio.load_system(q), E * 3.0)
    assert np.array_equal(sysio.load_system(p), E)


@pytest.mark.parametrize("bad", ["f4", "fortran", "shape7", "ndim3", "truncated", "notnpy"])
def test_rejects_other_arrays(tmp_path, bad):
    p = str(tmp_path / "bad.npy")
    a = np.zeros((2, 3, 4, 8))
    if bad == "f4":
        np.save(p, a.astype(np.float32))
    elif bad == "fortran":
        np.save(p, np.asfortranarray(a))
    elif bad == "shape7":
        np.save(p, np.zeros((2, 3, 4, 7)))
    elif bad == "ndim3":
        np.save(p, np.zeros((2, 3, 32)))
    elif bad == "truncated":
        np.save(p, a)
        with open(p, "r+b") as f:
            f.truncate(200)
    else:
        with open(p, "wb") as f:
            f.write(b"not a numpy file at all" * 4)
    with pytest.raises(L.AphpoisError):
        sysio.load_system(p)
