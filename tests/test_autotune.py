"""The tile autotuner's cache (aphros_amd/autotune.py): keyed by device,
precision, preconditioner and edge; a cached tile is reused only for the same
key.  CPU: the timing itself is replaced by a stub; GPU: a real tuning run
returns a tile the solver accepts and that solves like the default tile."""
import numpy as np
import pytest

from aphros_amd import _lib as L
from aphros_amd import autotune, synth
from aphros_amd.solver import Poisson


def test_cache_is_keyed_by_device(tmp_path, monkeypatch):
    monkeypatch.setenv("APHPOIS_TUNE_CACHE", str(tmp_path / "tiles.json"))
    calls = []
    dev = {"name": "GPU-A/256cu"}
    monkeypatch.setattr(autotune, "device_name", lambda device=0: dev["name"])

    def fake_tune(n, precision, precond=0, device=0):
        calls.append((dev["name"], n, precision, precond))
        return (64 if dev["name"].startswith("GPU-A") else 32, 0), {(64, 0): 1.0}

    monkeypatch.setattr(autotune, "tune_edge", fake_tune)
    assert autotune.tile_for(64, L.MIXED) == (64, 0)
    assert autotune.tile_for(64, L.MIXED) == (64, 0)      # cached
    assert len(calls) == 1
    assert autotune.tile_for(64, L.FP64) == (64, 0)       # other precision: tuned
    assert autotune.tile_for(64, L.MIXED, precond=L.PRECOND_MG) == (64, 0)
    assert len(calls) == 3
    dev["name"] = "GPU-B/80cu"                             # another device: tuned again
    assert autotune.tile_for(64, L.MIXED) == (32, 0)
    assert len(calls) == 4
    dev["name"] = "GPU-A/256cu"
    assert autotune.tile_for(64, L.MIXED) == (64, 0)      # the first device's entry kept
    assert len(calls) == 4
    assert autotune.tile_for(64, L.MIXED, retune=True) == (64, 0)
    assert len(calls) == 5


@pytest.mark.gpu
def test_tuned_tile_solves_like_default(tmp_path, monkeypatch):
    monkeypatch.setenv("APHPOIS_TUNE_CACHE", str(tmp_path / "tiles.json"))
    tile, times = autotune.tune_edge(32, L.MIXED, iters=40)
    print(f"tuned tile {tile}: {times}", flush=True)
    assert tile in times and all(t > 0 for t in times.values())
    e, _, _ = synth.system(32, 3, 0.1, 0.2, 1000.0, 21)
    E = e.reshape(32, 32, 32, 8)
    out = []
    for t in (None, tile):
        s = Poisson(E.shape[:3], method=L.PCG, precision=L.MIXED, tol=1e-8)
        if t:
            s.set_tile(*t)
        s.set_system(E)
        x, info = s.solve(None)
        s.close()
        assert info["converged"] == 1 and info["true_relres"] <= 1e-8
        out.append(x)
    # other tiles sum the dot products in another order: both solutions meet
    # the 1e-8 residual and agree to what that leaves open
    d = out[0] - out[1]
    assert np.abs(d - d.mean()).max() <= 1e-4 * np.abs(out[0]).max()
