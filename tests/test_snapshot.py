"""Snapshot I/O in the reference's ASCII format (CPU)."""
import os

import numpy as np

from fdps_sph_amd import snapshot as S

FIX = os.path.join(os.path.dirname(__file__), "golden", "tar_subset_2000.dat")


def test_read_reference_snapshot_subset():
    s = S.read_ascii(FIX)
    assert abs(s["time"] - 1.000777e4) < 1e-9
    assert len(s["mass"]) == 2000
    assert set(np.unique(s["tag"]).tolist()) == {0, 1}
    assert np.all(s["mass"] == float("437800000000000065536"))  # "%lf" of the reference's mass
    ic = S.to_ic(s)
    assert np.array_equal(ic["mat"], np.where(s["tag"] % 2 == 0, 0, 1))


def test_write_read_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    n = 50
    snap = {"time": 123.5, "id": np.arange(n), "tag": rng.integers(0, 2, n), "mass": rng.random(n) * 1e20,
            "pos": rng.normal(0, 1e6, (n, 3)), "vel": rng.normal(0, 1e3, (n, 3)), "dens": rng.random(n) * 3e3,
            "eng": rng.random(n) * 1e6, "pres": rng.random(n) * 1e9, "pot": -rng.random(n) * 1e7,
            "ent": rng.random(n) * 3e3, "temp": rng.random(n) * 1e3}
    p = tmp_path / "s.dat"
    S.write_ascii(str(p), snap)
    back = S.read_ascii(str(p))
    assert back["time"] == float("%e" % 123.5)
    assert np.array_equal(back["id"], snap["id"]) and np.array_equal(back["tag"], snap["tag"])
    for k in ("mass", "dens", "eng", "pres", "pot", "ent", "temp"):
        assert np.allclose(back[k], snap[k], rtol=0, atol=5e-7), k  # "%lf": 6 decimals
    assert np.allclose(back["pos"], snap["pos"], rtol=0, atol=5e-7)
    # the first line of the reference file is reproduced byte for byte
    with open(FIX) as f:
        first = [f.readline() for _ in range(3)]
    s = S.read_ascii(FIX)
    q = tmp_path / "t.dat"
    S.write_ascii(str(q), {k: (v[:1] if isinstance(v, np.ndarray) else v) for k, v in s.items()})
    with open(q) as f:
        assert f.readline() == first[0]
        f.readline()
        assert f.readline() == first[2]
