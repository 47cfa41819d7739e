"""Pins of the oracle that do not depend on the reference binary: closed
forms of the ZD and LJ pair terms, energy/force consistency, Newton's third
law, the self term, SHAKE residuals and the pair-list set on awkward boxes."""
import math

import numpy as np
import pytest

from j_presto_amd import systems
from oracle import Oracle

EEK = 332.06378


def two_ions(r, q1=1.0, q2=-1.0, rm=(1.2, 2.5), ep=(0.3, 0.04), L=40.0):
    """Two ions of different LJ types at separation r along x."""
    na = systems.ion_template("NA", 1, 1, 22.99, q1)
    cl = systems.ion_template("CL", 1, 2, 35.45, q2)
    types = [systems.AtomType("A", rm[0], ep[0]), systems.AtomType("B", rm[1], ep[1])]
    x = np.array([[10.0, 10.0, 10.0], [10.0 + r, 10.0, 10.0]])
    return systems.System("pair", [na, cl], types, np.array([L, L, L]), x, np.zeros_like(x))


def pair_energy_force(r, **kw):
    s = two_ions(r, **kw)
    o = Oracle(s)
    assert o.celprp(s.x) == 0
    g, ee, ev = o.nonbond(s.x)
    return s, o, g[0], ee[0], ev[0]


@pytest.mark.parametrize("r", [2.0, 3.3, 7.5, 11.9])
def test_zd_pair_closed_form(r):
    """Zero-dipole (alpha=0): E = qq (1/r + r^2/(2 rc^3) - 3/(2 rc)),
    dE/dr = qq (-1/r^2 + r/rc^3), plus the self term -zcore/2 ccoef sum q^2."""
    rc = 12.0
    s, o, g, ee, ev = pair_energy_force(r, ep=(0.0, 0.0))
    qq = EEK * 1.0 * -1.0
    e_pair = qq * (1 / r + r * r / (2 * rc ** 3) - 3 / (2 * rc))
    e_self = -(1.5 / rc) * 0.5 * EEK * (1.0 + 1.0)
    assert math.isclose(ee, e_pair + e_self, rel_tol=1e-13)
    dedr = qq * (-1 / r ** 2 + r / rc ** 3)
    # atom 0 sits at smaller x: dE/dx0 = -dE/dr
    assert math.isclose(g[0, 0], -dedr, rel_tol=1e-12)
    assert math.isclose(g[1, 0], dedr, rel_tol=1e-12)
    assert abs(g[0, 1]) == 0 and abs(g[0, 2]) == 0


def test_zd_vanishes_at_and_beyond_cutoff():
    rc = 12.0
    e_self = -(1.5 / rc) * 0.5 * EEK * 2.0
    for r in (12.0, 12.5, 13.4):
        s, o, g, ee, ev = pair_energy_force(r, ep=(0.0, 0.0))
        if r == 12.0:
            # r == rc is inside (co = sign(0.5,0)+0.5 = 1) and E(rc) = F(rc) = 0
            assert abs(ee - e_self) < 1e-12 and np.abs(g).max() < 1e-12
        else:
            assert ee == e_self and np.abs(g).max() == 0.0


@pytest.mark.parametrize("r", [3.0, 3.7, 5.0])
def test_lj_amber_closed_form(r):
    """AMBER combination: C6 = 2 sqrt(e1 e2) (R1+R2)^6, C12 = sqrt(e1 e2)
    (R1+R2)^12; minimum -sqrt(e1 e2) at r = R1+R2, truncated at rc."""
    rm, ep = (1.2, 2.5), (0.3, 0.04)
    s, o, g, ee, ev = pair_energy_force(r, q1=0.0, q2=0.0, rm=rm, ep=ep)
    e = math.sqrt(ep[0] * ep[1])
    R = rm[0] + rm[1]
    assert math.isclose(ev, e * ((R / r) ** 12 - 2 * (R / r) ** 6), rel_tol=1e-12)
    dedr = e * (-12 * R ** 12 / r ** 13 + 12 * R ** 6 / r ** 7)
    assert math.isclose(g[1, 0], dedr, rel_tol=1e-10, abs_tol=1e-12)
    if r == 3.7:
        assert math.isclose(ev, -e, rel_tol=1e-12) and abs(g[1, 0]) < 1e-12


def test_lj_table_and_zero_vdw():
    s = systems.water_box(100, 24.5, 3, n_ion_pairs=2)
    o = Oracle(s)
    c6, c12 = o.ljtable()
    R = np.array([t.rmin2 for t in s.types])
    E = np.array([t.eps for t in s.types])
    assert np.allclose(c6, 2 * np.sqrt(np.outer(E, E)) * np.add.outer(R, R) ** 6, rtol=1e-14)
    assert np.allclose(c12, np.sqrt(np.outer(E, E)) * np.add.outer(R, R) ** 12, rtol=1e-14)
    assert np.all(c6[1] == 0) and np.all(c12[1] == 0)   # TIP3P H has no LJ


def test_energy_force_finite_difference():
    s = systems.water_box(500, 24.8, 7)
    o = Oracle(s)
    x = s.x.copy()
    o.celprp(x)
    g, _, _ = o.nonbond(x)
    rng = np.random.default_rng(0)
    h = 1e-5
    checked = 0
    for a in rng.permutation(s.n)[:40]:
        d = x - x[a]
        d -= s.box * np.round(d / s.box)
        r = np.sqrt((d * d).sum(1))
        if np.any(np.abs(r - 12.0) < 1e-3):   # LJ is truncated, not shifted, at rc
            continue
        for k in range(3):
            xp, xm = x.copy(), x.copy()
            xp[a, k] += h
            xm[a, k] -= h
            _, ep, vp = o.nonbond(xp)
            _, em, vm = o.nonbond(xm)
            fd = ((ep + vp) - (em + vm)).sum() / (2 * h)
            assert abs(fd - g[0, a, k]) < 2e-5 * max(1.0, abs(g[0, a, k])), (a, k, fd, g[0, a, k])
        checked += 1
    assert checked >= 20


def test_newton_third_law_and_self_term():
    s = systems.water_box(500, 24.8, 7, n_ion_pairs=3)
    o = Oracle(s)
    o.celprp(s.x)
    g, _, _ = o.nonbond(s.x)
    assert np.abs(g[0].sum(0)).max() < 1e-9 * np.abs(g).max() * s.n
    # neutral, fully excluded waters contribute nothing to dself; ions contribute -zcore/2 ccoef q^2
    ds = o.dself()[0]
    assert math.isclose(ds, -(1.5 / 12.0) * 0.5 * EEK * 6.0, rel_tol=1e-12)


def test_shake_residuals_and_momentum():
    s = systems.chain_in_water(40, 25.0, 5)
    o = Oracle(s)
    rng = np.random.default_rng(1)
    x0 = s.x.copy()
    x1 = x0 + rng.normal(scale=0.02, size=x0.shape)
    x1c = x1.copy()
    assert o.shake(x1c, x0) == 0
    for atoms, d2 in s.shake_groups():
        pairs = [(0, 1)] if len(atoms) == 2 else [(0, 1), (1, 2), (2, 0)]
        for (p, q), t in zip(pairs, d2):
            d = x1c[atoms[p]] - x1c[atoms[q]]
            d -= s.box * np.round(d / s.box)
            # 3-atom groups converge to |B| < tol*d^2; 2-atom groups are exact
            assert abs(d @ d - t) <= (1e-6 * t if len(atoms) == 3 else 1e-12 * t)
    m = s.mass[:, None]
    assert np.abs((m * (x1c - x1)).sum(0)).max() < 1e-12


def test_verlet_free_flight():
    """No forces, no constraints: x(t+dt) = x + dt v exactly, v unchanged."""
    s = two_ions(20.0, q1=0.0, q2=0.0, ep=(0.0, 0.0))
    o = Oracle(s)
    x, v = s.x.copy(), np.array([[0.01, -0.02, 0.005], [0.0, 0.001, -0.003]])
    x0, v0 = x.copy(), v.copy()
    assert o.verlet_step(x, v, np.zeros_like(x)) == 0
    assert np.array_equal(v, v0) and np.array_equal(x, x0 + 2.0 * v0)


@pytest.mark.parametrize("L", [24.05, 27.0, 40.5, 33.3])
def test_list_set_awkward_boxes(L):
    """Boxes just above 2*rc, and edges that are multiples of rlim = 13.5 A
    (md_run's phantom-cell case): the cell list is still the exact set."""
    n = int(systems.WATER_PER_A3 * L ** 3 * 0.9)
    s = systems.water_box(n, L, 11)
    o = Oracle(s)
    assert o.celprp(s.x) == 0
    cell, ptr, idx = o.get_list()
    bp, bi = o.list_brute(s.x)
    assert np.array_equal(ptr, bp) and np.array_equal(idx, bi)
    c = o.constants()
    assert cell.min() >= c["mx1"] and cell.max() <= c["mx2"]


def test_exclusion_out_of_range_is_ier99():
    """An excluded pair farther apart than rlim is not in the list: cutoff_PB
    finds fewer pairs to remove and returns ier = 99 (cutoff.f90:90-96)."""
    s = systems.water_box(100, 24.5, 4)
    x = s.x.copy()
    x[1] += np.array([12.0, 12.0, 0.0])   # pull H1 of water 0 ~17 A from its O
    o = Oracle(s)
    assert o.celprp(x) == 99


def test_ekin_closed_form():
    """caltmp's kinetic energy is (1/2) sum m v^2 in kcal/mol: with m in g/mol
    and v in A/fs (1 A/fs = 1e5 m/s), SI arithmetic gives the same value up
    to caltmp's factor 0.0005, a single-precision literal (caltmp.f90:43):
    K = (1/2) sum m v^2 * float32(0.0005) / 0.0005."""
    from j_presto_amd import systems
    s = systems.water_box(8, 12.5, 3)
    o = Oracle(s)
    rng = np.random.default_rng(4)
    v = rng.normal(scale=0.01, size=(s.n, 3))
    m_kg = s.mass * 1e-3                       # kg/mol
    v_si = v * 1e5                             # m/s
    lit = float(np.float32(0.0005)) / 0.0005   # 1 + 4.75e-8
    k_si = 0.5 * (m_kg[:, None] * v_si * v_si).sum() / 4184.0   # kcal/mol
    assert abs(o.ekin(v) - k_si * lit) <= 1e-12 * k_si
    assert abs(o.ekin(v) / k_si - 1.0) < 5e-8
    # one atom moving alone
    v1 = np.zeros_like(v)
    v1[4] = [3e-3, 0.0, 0.0]
    assert abs(o.ekin(v1) - 0.5 * s.mass[4] * 1e-3 * (300.0 ** 2) / 4184.0 * lit) < 1e-15
