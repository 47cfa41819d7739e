"""The oracle's centre-of-mass stopping (STOPCE, md_run's remotv) and the
initial step of a run without a restart (RESTAR= NO, INITIA= ZERO: staver)
bit for bit against the compiled reference md_run (STOP_RUNS of
tests/golden/make_golden.py --stop-only on chain_tors, chain_lmd_tors and
chain_droplet_cmm), and remotv's three modes against what they must leave:
no momentum of the stopped chain (TRAN), no angular momentum about its
centre of mass (ROTA), neither (BOTH), every other atom untouched.  md_run's
own STOPCE= TRAN reads an unset angular velocity (remotv.f90:54-69,
122-127), so TRAN is pinned by these invariants alone."""
import numpy as np
import pytest

from oracle import Oracle

STOP_NAMES = ["chain_tors", "chain_lmd_tors", "chain_droplet_cmm"]
# tag -> (METHOD, STOPCE, RESTAR, steps, TEMPCO), as the fixtures were made
STOP_RUNS = {"mb": ("MICR", "BOTH", "YES", 12, "NO"), "mr": ("MICR", "ROTA", "YES", 12, "NO"),
             "cb": ("CANO", "BOTH", "YES", 12, "NO"), "cy": ("CANO", "ROTA", "YES", 12, "YES"),
             "s0": ("MICR", "NO", "NO", 12, "NO"), "sc": ("CANO", "BOTH", "NO", 12, "NO")}


def oracle_stop_run(s, tag, cmm):
    method, stopce, restar, n, tempco = STOP_RUNS[tag]
    o = Oracle(s)
    if cmm:
        assert o.cmm_setup(s.x) == 0
    assert o.set_stopce(stopce, s.chain_end(1), s.first_kind_atoms()) == 0
    x, v = s.x.copy(), s.v.copy()
    if restar == "NO":   # INITIA= ZERO, then staver
        v[:] = 0.0
        assert o.start_step(x, v) == 0
    if method == "MICR":   # md_run's restart set-up step leaves x and v as they were
        ier = o.cmm_run_nve(x, v, n, 10, 5) if cmm else o.run_nve(x, v, n, 10)
    else:
        grp = s.temperature_groups() if tempco == "YES" else None
        assert not cmm
        ier, _ = o.run_cano(x, v, n, 10, 300.0, grp=grp, inirst=restar == "YES")
    return ier, x, v


@pytest.mark.parametrize("name", STOP_NAMES)
@pytest.mark.parametrize("tag", list(STOP_RUNS))
def test_stop_runs_vs_reference(golden, name, tag):
    s, f = golden(name)
    cmm = name == "chain_droplet_cmm"
    if cmm and STOP_RUNS[tag][0] == "CANO":
        pytest.skip("the oracle's canonical loop runs the ZD path only; CMM canonical runs are GPU tests")
    ier, x, v = oracle_stop_run(s, tag, cmm)
    assert ier == 0
    dx, dv = np.abs(x - f[f"stop_{tag}_x"]).max(), np.abs(v - f[f"stop_{tag}_v"]).max()
    print(name, tag, "max |dx|", dx, "max |dv|", dv)
    assert dx == 0.0 and dv == 0.0


def moments(s, x, v, n):
    m = s.mass[:n, None]
    c = (m * x[:n]).sum(0) / m.sum()
    p = (m * v[:n]).sum(0)
    L = (m * np.cross(x[:n] - c, v[:n])).sum(0)
    return p, L


@pytest.mark.parametrize("mode", ["TRAN", "ROTA", "BOTH"])
def test_remotv_invariants(golden, mode):
    s, _ = golden("chain_tors")
    n = s.chain_end(1)
    o = Oracle(s)
    assert o.set_stopce(mode, n) == 0
    rng = np.random.default_rng(4)
    v0 = s.v + rng.normal(scale=0.01, size=s.v.shape) + np.array([0.003, -0.002, 0.001])
    v = v0.copy()
    assert o.remotv(s.x, v) == 0
    p0, L0 = moments(s, s.x, v0, n)
    p, L = moments(s, s.x, v, n)
    print(mode, "p", np.abs(p0).max(), "->", np.abs(p).max(), "L", np.abs(L0).max(), "->", np.abs(L).max())
    assert np.array_equal(v[n:], v0[n:])
    dv = v[:n] - v0[:n]
    if mode in ("TRAN", "BOTH"):
        assert np.abs(p).max() < 1e-12 * np.abs(p0).max()
    if mode == "TRAN":   # a uniform shift: the angular momentum changes only by -M c x v_cm
        assert np.abs(dv - dv[0]).max() < 1e-15
    if mode in ("ROTA", "BOTH"):
        assert np.abs(L).max() < 1e-10 * np.abs(L0).max()
    if mode == "ROTA":   # a rigid rotation about the centre of mass: momentum kept
        assert np.abs(p - p0).max() < 1e-12 * np.abs(p0).max()
        m = s.mass[:n, None]
        c = (m * s.x[:n]).sum(0) / m.sum()
        w = np.linalg.lstsq(np.cross(np.eye(3)[None, :, :], (s.x[:n] - c)[:, None, :]).reshape(-1, 3) * -1,
                            dv.reshape(-1), rcond=None)[0]
        assert np.abs(np.cross(w, s.x[:n] - c) - dv).max() < 1e-12


def test_remotv_errors(golden):
    """A single stopped atom at the origin has an exactly zero inertia tensor
    (md_run's ier -2, -3 here: md_run's SHAKE uses -2), and v stays as it
    was; TRAN alone needs no inverse and stops the atom."""
    s, _ = golden("chain_tors")
    o = Oracle(s)
    xp = s.x.copy()
    xp[0] = 0.0
    v = s.v.copy()
    assert o.set_stopce("BOTH", 1) == 0
    assert o.remotv(xp, v) == -3
    assert o.set_stopce("ROTA", 1) == 0
    assert o.remotv(xp, v) == -3
    assert np.array_equal(v, s.v)
    assert o.set_stopce("TRAN", 1) == 0
    assert o.remotv(xp, v) == 0
    assert np.abs(v[0]).max() < 1e-14 * np.abs(s.v[0]).max() and np.array_equal(v[1:], s.v[1:])


def test_stop_dof(golden):
    """STOPCE's degrees of freedom: 6 (BOTH) or 3 off the one group; under
    TEMPCO= YES split over the first molecule kind's groups in proportion to
    their degrees of freedom (md.f90:183-215)."""
    s, _ = golden("chain_lmd_tors")
    o = Oracle(s)
    base1 = o.group_dof(None)
    # the generator's molecules sit in one cluster each: split the chain over
    # two groups by hand, as per-atom clusters of md_run's input would
    grp = s.temperature_groups() + 1
    grp[:s.first_kind_atoms()] = np.arange(s.first_kind_atoms()) % 2
    baseg = o.group_dof(grp)
    assert o.set_stopce("BOTH", s.chain_end(1), s.first_kind_atoms()) == 0
    assert np.array_equal(o.group_dof(None), base1 - 6.0)
    d = o.group_dof(grp)
    ifrg = int(grp[:s.first_kind_atoms()].max()) + 1
    assert ifrg >= 2
    assert np.allclose(baseg[:ifrg] - d[:ifrg], 6.0 * baseg[:ifrg] / baseg[:ifrg].sum(), rtol=0, atol=1e-12)
    assert np.array_equal(d[ifrg:], baseg[ifrg:])
    assert abs((baseg - d).sum() - 6.0) < 1e-12
