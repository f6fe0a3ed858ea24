"""CPU check of the plane-sampled L1 comparison of idefix_amd.accuracy (used for runs that
do not fit one process pair): identical files give 0, a known offset gives that offset."""
import numpy as np
import pytest

from idefix_amd import accuracy


def _save(path, planes, t=1.0):
    np.savez(path, k=np.array([5, 15]), planes=planes, t=t, cycles=10, wall_s=1.0,
             zone_updates_per_s=1e6, divB=1e-14)


def test_compare_planes(tmp_path):
    rng = np.random.default_rng(0)
    p = rng.normal(size=(2, 8, 6, 7))
    _save(tmp_path / "a.npz", p)
    _save(tmp_path / "b.npz", p)
    rep = accuracy.compare_planes(tmp_path / "a.npz", tmp_path / "b.npz")
    assert rep["L1"] == 0.0
    q = p.copy()
    q[:, 3] += 0.25
    _save(tmp_path / "c.npz", q)
    rep = accuracy.compare_planes(tmp_path / "a.npz", tmp_path / "c.npz")
    assert rep["L1_per_variable"][3] == pytest.approx(0.25, rel=1e-12)
    assert rep["L1"] == pytest.approx(0.25 / 8, rel=1e-12)
    _save(tmp_path / "d.npz", p, t=0.5)
    with pytest.raises(AssertionError):
        accuracy.compare_planes(tmp_path / "a.npz", tmp_path / "d.npz")
