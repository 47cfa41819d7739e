"""Synthetic systems and md_run file formats."""
import os

import numpy as np

from j_presto_amd import refio, systems


def test_water_geometry_and_contacts():
    s = systems.water_box(300, 24.5, 2)
    x = s.x.reshape(-1, 3, 3)
    L = s.box
    def dist(a, b):
        d = a - b
        d -= L * np.round(d / L)
        return np.sqrt((d * d).sum(-1))
    assert np.allclose(dist(x[:, 0], x[:, 1]), systems.R_OH, atol=1e-12)
    assert np.allclose(dist(x[:, 0], x[:, 2]), systems.R_OH, atol=1e-12)
    assert np.allclose(dist(x[:, 1], x[:, 2]), systems.R_HH, atol=1e-12)
    from scipy.spatial import cKDTree
    t = cKDTree(np.mod(s.x, L), boxsize=L)
    pairs = t.query_pairs(1.6, output_type="ndarray")
    assert np.all(pairs[:, 0] // 3 == pairs[:, 1] // 3)
    assert abs(s.charge.sum()) < 1e-12


def test_velocities_respect_constraints_and_temperature():
    s = systems.chain_in_water(60, 25.0, 3)
    x, v, L = s.x, s.v, s.box
    for atoms, d2 in s.shake_groups():
        pairs = [(0, 1)] if len(atoms) == 2 else [(0, 1), (1, 2), (2, 0)]
        for p, q in pairs:
            d = x[atoms[p]] - x[atoms[q]]
            d -= L * np.round(d / L)
            assert abs(d @ (v[atoms[p]] - v[atoms[q]])) < 1e-12
    T = (s.mass[:, None] * v * v).sum() * 1e7 / (systems.GASCST * (s.ndf() - 3))
    assert abs(T - 300.0) < 1e-9
    assert np.abs((s.mass[:, None] * v).sum(0)).max() < 1e-12


def test_restart_roundtrip(tmp_path):
    s = systems.water_box(50, 24.5, 1)
    p = os.path.join(tmp_path, "r.rst")
    refio.write_restart(p, s.x, s.v, loop=7, time_fs=14.0, lam=0.3)
    r = refio.read_restart(p)
    assert r["n"] == s.n and r["loop"] == 7 and r["lam"] == 0.3
    assert np.array_equal(r["x"], s.x) and np.array_equal(r["v"], s.v)


def test_tpl_lines_fit_md_run_reader(tmp_path):
    s = systems.chain_in_water(40, 25.0, 5)
    files = refio.write_inputs(str(tmp_path), s)
    lines = open(os.path.join(tmp_path, files["TOPOLO"])).read().splitlines()
    assert max(len(l) for l in lines) < 80          # md_run reads a80
    assert lines[0].split()[1] == str(s.n)


def test_workload_sizes():
    assert systems.workload(0).n == 6000
    s2_L = (33333 / systems.WATER_PER_A3) ** (1 / 3)
    assert abs(s2_L - 100.0) < 0.05
