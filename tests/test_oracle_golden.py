"""The CPU oracle against the compiled reference md_run (golden fixtures
made by tests/golden/make_golden.py): bit-exact forces, energies, pair-list
sets and NVE trajectories."""
import numpy as np
import pytest

from conftest import GOLDEN_NAMES, ge_params, oracle_monitor
from oracle import Oracle


@pytest.mark.parametrize("name", GOLDEN_NAMES)
def test_forces_and_energies_bit_exact(golden, name):
    s, f = golden(name)
    o = Oracle(s)
    assert o.celprp(s.x) == 0
    # dnergy order: bonds and angles, then cmmdo, into one zeroed gradient
    g, et = o.bonded(s.x)
    g, ee, ev = o.nonbond(s.x, g)
    gt = o.combine(g)
    assert np.array_equal(gt, f["force_grad"]), np.abs(gt - f["force_grad"]).max()
    # dnergy slots (main.f90:157-159): 2 BOND, 3 ANGLE, 4 TORSION, 5 IMPROPER, 6 1-4 VDW,
    # 7 1-4 ELEC, 8 VDW15, 9 ELE15; unscaled class sums
    assert ee.sum() == f["force_epot"][8]
    assert ev.sum() == f["force_epot"][7]
    for k in range(6):
        assert et[:, k].sum() == f["force_epot"][1 + k], k


@pytest.mark.parametrize("name", GOLDEN_NAMES)
def test_cell_list_is_bruteforce_set(golden, name):
    """The list is {j : FP64 minimum-image d^2 <= rlim^2} minus the excluded
    pairs, restricted to the cell pairs md_run's periodic neighbour-cell
    table holds (restated independently in conftest.md_cell_windows).  On
    the 42 A box that table drops pairs inside the cutoff; md_run's own
    forces there (the fixture) agree with the oracle's bit for bit."""
    from conftest import admissible, md_cell_windows
    s, f = golden(name)
    o = Oracle(s)
    assert o.celprp(s.x) == 0
    cell, ptr, idx = o.get_list()
    bp, bi = o.list_brute(s.x)
    win = md_cell_windows(s.box, int(o.constants()["nv"]))
    rows = np.repeat(np.arange(s.n), np.diff(bp))
    keep = admissible(win, cell[rows], cell[bi - 1])
    dropped = int((~keep).sum())
    if name == "water_42":
        d = s.x[rows[~keep]] - s.x[bi[~keep] - 1]
        d -= s.box * np.round(d / s.box)
        r = np.sqrt((d * d).sum(1))
        print(f"{name}: md_run's cell table drops {dropped} of {len(bi)} listed pairs, "
              f"{int((r <= s.cutoff).sum())} of them within the cutoff (closest {r.min():.2f} A)")
        assert (r <= s.cutoff).sum() > 0
    else:
        assert dropped == 0
    cnt = np.bincount(rows[keep], minlength=s.n)
    assert np.array_equal(ptr, np.concatenate([[0], np.cumsum(cnt)]))
    assert np.array_equal(idx, bi[keep])


@pytest.mark.parametrize("name", [n for n in GOLDEN_NAMES if n not in ("ions_3k", "water_54")])
def test_nve_trajectory_bit_exact(golden, name):
    """md_run METHOD= MICR from a restart: K steps of force, kick, drift,
    SHAKE, v=(x-x0)/dt, list rebuilt after every 10th step (md.f90:519-694;
    inirst restores x and v after its check step, md.f90:1241-1256)."""
    s, f = golden(name)
    o = Oracle(s)
    x, v = s.x.copy(), s.v.copy()
    assert o.celprp(x) == 0
    for idloop in range(1, int(f["traj_nsteps"]) + 1):
        g, et = o.bonded(x)
        g, ee, ev = o.nonbond(x, g)
        assert o.verlet_step(x, v, o.combine(g)) == 0
        if idloop % 10 == 0:
            assert o.celprp(x) == 0
    assert np.array_equal(x, f["traj_x"])
    assert np.array_equal(v, f["traj_v"])


def test_potential_energy_series(golden):
    """NAMETO prints E_pot of every step with 8 significant digits."""
    s, f = golden("water_1500")
    o = Oracle(s)
    x, v = s.x.copy(), s.v.copy()
    o.celprp(x)
    ep = []
    for idloop in range(1, int(f["traj_nsteps"]) + 1):
        g, ee, ev = o.nonbond(x)
        ep.append(ee.sum() + ev.sum())
        o.verlet_step(x, v, o.combine(g))
        if idloop % 10 == 0:
            o.celprp(x)
    ref = f["traj_epot"][:, 0]
    # E15.7 keeps 7 significant digits: compare the same decimal rounding
    assert [float(f"{e:.6e}") for e in ep] == list(ref)


@pytest.mark.parametrize("name", ["water_1500", "chain_3k", "chain_lmd", "methyl_water"])
def test_cano_trajectory_bit_exact(golden, name):
    """md_run METHOD= CANO (verlethe.f90) from a restart, TEMPCO= NO (one
    temperature) or YES (one per molecule kind and cluster): velocity scaling
    to 300 K, then the SHAKE/temperature rounds, list rebuilt after every
    10th step."""
    s, f = golden(name)
    o = Oracle(s)
    grp = None if str(f["cano_tempco"]) == "NO" else s.temperature_groups()
    x, v = s.x.copy(), s.v.copy()
    assert o.celprp(x) == 0
    for idloop in range(1, int(f["cano_nsteps"]) + 1):
        g, et = o.bonded(x)
        g, ee, ev = o.nonbond(x, g)
        ier, vb = o.cano_step(x, v, o.combine(g), 300.0, 3, 1e-3, grp)
        assert ier == 0
        if idloop % 10 == 0:
            assert o.celprp(x) == 0
    assert np.array_equal(x, f["cano_x"])
    assert np.array_equal(v, f["cano_v"])


def test_cano_groups_and_dof():
    """Temperature groups and their degrees of freedom in closed form: rigid
    water has 6 per molecule; a chain in water splits into the chain's kind
    and the water's kind (TEMPCO= YES), each with 3 per atom minus its own
    SHAKE constraints (one per X-H group on the chain, three per water)."""
    from j_presto_amd import systems
    s = systems.water_box(60, 14.0, 3)
    o = Oracle(s)
    assert o.group_dof()[0] == 6 * 60
    assert s.temperature_groups().max() == 0
    c = systems.chain_in_water(n_chain=40, L=20.0, seed=5)
    grp = c.temperature_groups()
    nch = int((grp == 0).sum())
    nw = (c.n - nch) // 3
    n_xh = sum(1 for atoms, _ in c.shake_groups() if len(atoms) == 2)
    dof = Oracle(c).group_dof(grp)
    assert list(dof) == [3.0 * nch - n_xh, 6.0 * nw]


def printed(a):
    """md_run's e15.7 rounding: 7 significant digits."""
    return [float(f"{e:.6e}") for e in np.ravel(a)]


@pytest.mark.parametrize("name,method", [("water_1500", "micr"), ("chain_lmd", "micr"),
                                         ("water_1500", "cano"), ("chain_lmd", "cano")])
def test_monitored_energies_vs_reference(golden, name, method):
    """caltmp's kinetic energy, E_pot and E_tot at md_run's OUTLOG steps
    (monmd.f90 outtlg) equal the oracle's to every printed digit, and the
    printed temperature is K over md_run's degrees of freedom (3 per atom
    minus one per SHAKE constraint, md.f90:114-118)."""
    s, f = golden(name)
    mon = f[f"mon_{method}"]
    o = Oracle(s)
    cano = None
    if method == "cano":
        cano = (300.0, None if str(f["cano_tempco"]) == "NO" else s.temperature_groups())
    nsteps = int(f["traj_nsteps"] if method == "micr" else f["cano_nsteps"])
    rows = oracle_monitor(o, s, s.x.copy(), s.v.copy(), nsteps, 5, cano)
    ref = mon[mon[:, 0] <= nsteps]
    assert np.array_equal(rows[:, 0], ref[:, 0])
    assert printed(rows[:, 3]) == list(ref[:, 4])      # KINETIC
    assert printed(rows[:, 2]) == list(ref[:, 2])      # POTENTIAL
    assert printed(rows[:, 1]) == list(ref[:, 1])      # TOTAL
    # TEMPERATURE (f15.7) = sum m v^2 / (R ndf) in md_run's units (caltmp.f90:40-43,
    # whose kinetic-energy factor 0.0005 is a single-precision literal)
    temp = rows[:, 3] * 4.184 / float(np.float32(0.0005)) / (8.31451 * s.ndf())
    assert np.allclose(temp, ref[:, 3], rtol=0, atol=1e-6)


def test_oracle_run_nve_equals_reference_trajectory(golden):
    """Oracle.run_nve (the oracle's own MICR loop: dnergy, verlet, rebuild
    after every `update`-th step) reproduces md_run's trajectory bit for bit."""
    for name in ("water_1500", "chain_bonded"):
        s, f = golden(name)
        o = Oracle(s)
        x, v = s.x.copy(), s.v.copy()
        assert o.run_nve(x, v, int(f["traj_nsteps"]), 10) == 0
        assert np.array_equal(x, f["traj_x"]) and np.array_equal(v, f["traj_v"]), name


def test_cano_shake_retry_equals_reference(golden):
    """METHOD= CANO with LIMSHK= 3 (SHAKE fails at every full 2 fs step):
    Oracle.run_cano's verlethe retry (undo, halve dt, redo the forces, double
    back after two half substeps) and md_run's restart set-up step reproduce
    md_run's trajectory bit for bit.  md_run printed one provision more than
    the oracle counts: finmd repeats a step from the final state for its
    report and discards it (md.f90:905-955).  Without the restart set-up step
    (inirst's velocities from verlethe's velback) the trajectory differs."""
    s, f = golden("water_1500")
    o = Oracle(s, shake_maxit=3)
    x, v = s.x.copy(), s.v.copy()
    ier, nprov = o.run_cano(x, v, int(f["retry_loop"]), 10, 300.0, 3, 1e-3)
    assert ier == 0
    assert nprov + 1 == int(f["retry_provisions"])
    assert np.array_equal(x, f["retry_x"]) and np.array_equal(v, f["retry_v"])
    o = Oracle(s, shake_maxit=3)
    x, v = s.x.copy(), s.v.copy()
    ier, nprov = o.run_cano(x, v, int(f["retry_loop"]), 10, 300.0, 3, 1e-3, inirst=False)
    assert ier == 0 and nprov + 2 == int(f["retry_provisions"])
    assert np.abs(x - f["retry_x"]).max() > 1e-3


@pytest.mark.parametrize("tag", ["win", "lo", "hi"])
def test_ge_trajectory_equals_reference(golden, tag):
    """METHOD= GE (multicanonical, verlethe.f90:196-233): the forces carry
    the weight of the class-1 energy, from the window polynomials ("win":
    the run's reduced energy falls from 0.87 to about 0.2 through all three
    windows), alphalw below celw ("lo") or alphaup above ceup ("hi").
    Oracle.run_cano with the parameters reproduces md_run bit for bit; the
    same run with the two constant weights exchanged does not."""
    s, f = golden("water_1500")
    ge = ge_params(f, tag)
    n = int(f[f"ge_{tag}_loop"])
    o = Oracle(s)
    x, v = s.x.copy(), s.v.copy()
    assert o.run_cano(x, v, n, 10, 300.0, 3, 1e-3, ge=ge) == (0, 0)
    assert np.array_equal(x, f[f"ge_{tag}_x"]) and np.array_equal(v, f[f"ge_{tag}_v"])
    if tag != "win":
        bad = dict(ge, alphalw=ge["alphaup"], alphaup=ge["alphalw"])
        o = Oracle(s)
        x, v = s.x.copy(), s.v.copy()
        assert o.run_cano(x, v, n, 10, 300.0, 3, 1e-3, ge=bad) == (0, 0)
        assert np.abs(x - f[f"ge_{tag}_x"]).max() > 1e-6


@pytest.mark.parametrize("tag", ["a", "b", "c", "d"])
def test_alsd_trajectory_equals_reference(golden, tag):
    """Adaptive lambda square dynamics (METHOD= GE under GEPSME= LMD,
    verlethe.f90:235-290): lambda moves under -2 E1 lambda - E2 - dew beside
    the atoms, its kinetic term in the temperature.  Runs inside the windows
    ("a", "d"), below celw ("b": lower wall) and above ceup ("c": upper
    wall), with lambda's temperature control ("a", "b") and without ("c",
    "d"), one temperature group ("a", "c") and TEMPCO= YES ("b", "d").
    Oracle.run_alsd reproduces md_run's x, v, lambda and lambda velocity bit
    for bit, and lambda after every step to NAMETO's printed digits; the same
    run with the walls exchanged, or with the temperature-control flag
    flipped, does not."""
    from conftest import alsd_params
    s, f = golden("chain_lmd")
    al, tempco = alsd_params(s, f, tag)
    grp = s.temperature_groups() if tempco else None
    n = int(f[f"alsd_{tag}_loop"])
    o = Oracle(s)
    assert o.lam == s.lam
    x, v = s.x.copy(), s.v.copy()
    ier, lam, lamv, lam_log, nprov = o.run_alsd(x, v, 0.0, n, al, grp=grp)
    assert (ier, nprov) == (0, 0)
    assert np.array_equal(x, f[f"alsd_{tag}_x"]) and np.array_equal(v, f[f"alsd_{tag}_v"])
    assert lam == float(f[f"alsd_{tag}_lam"]) and lamv == float(f[f"alsd_{tag}_lamv"])
    ene = f[f"alsd_{tag}_ene"]
    assert np.allclose(lam_log, ene[:, 0], rtol=1e-7, atol=0)
    for bad in (dict(al, lower=al["upper"], upper=al["lower"]), dict(al, tctl=not al["tctl"])):
        if tag in ("a", "d") and bad["lower"] != al["lower"]:
            continue   # inside the windows the walls do not act
        o = Oracle(s)
        x, v = s.x.copy(), s.v.copy()
        ier, lam2, _, _, _ = o.run_alsd(x, v, 0.0, n, bad, grp=grp)
        assert ier == 0
        assert abs(lam2 - lam) > 1e-9 or np.abs(x - f[f"alsd_{tag}_x"]).max() > 1e-9


def test_alsd_monitored_temperature(golden):
    """md_run's monitoring of ALSD run "a" (monmd.f90:112 via caltmp with
    lambda_vt = (lambda_v before + after the step)/2, md.f90:618-633): the
    printed kinetic energy and temperature take lambda's term lm lvt^2 beside
    the atoms' (v(t-dt/2)+v(t+dt/2))/2, and the temperature counts lambda's
    degree of freedom (md.f90:114-123).  The term is far below K's printed
    digits; the temperature's seven decimals see both."""
    from conftest import alsd_params
    s, f = golden("chain_lmd")
    al, tempco = alsd_params(s, f, "a")
    assert al["tctl"]
    mon = f["alsd_a_mon"]
    o = Oracle(s)
    x, v = s.x.copy(), s.v.copy()
    ier, _, lv, _, _ = o.run_alsd(x, v, 0.0, 4, al)
    assert ier == 0
    vb, lvb = v.copy(), lv
    ier, _, lv, _, _ = o.run_alsd(x, v, lv, 1, al, inirst=False)
    assert ier == 0
    w = 0.5 * (vb + v)
    mw2 = float((s.mass[:, None] * w * w).sum())
    lvt = 0.5 * (lvb + lv)
    lke = al["lmass"] * lvt * lvt
    dof = float(o.group_dof(None)[0])
    row = mon[mon[:, 0] == 5][0]
    k = (mw2 + lke) * 1.0e7 / 4.184 * float(np.float32(0.0005))
    assert abs(k - row[4]) <= 5e-7 * abs(row[4])   # seven significant digits (E15.7)
    temp = lambda e, n: e * 1.0e7 / (8.31451 * n)
    assert abs(temp(mw2 + lke, dof + 1.0) - row[3]) <= 6e-8
    assert abs(temp(mw2, dof + 1.0) - row[3]) > 1e-6
    assert abs(temp(mw2 + lke, dof) - row[3]) > 1e-3


@pytest.mark.parametrize("name", ["water_1500", "chain_lmd"])
@pytest.mark.parametrize("tag", ["a", "b"])
def test_heating_loop_bit_exact(golden, name, tag):
    """md_run's heating loop (HEATLO= 10, STARTT= 100 K, SETTEM= 300 K): the
    target temperature of step idloop <= 10 is 100 + (200/10) idloop, the
    restart set-up step taking the restart's loop count (0 or 4); the final
    x and v bit-exact to md_run's, and a run without the ramp differs."""
    s, f = golden(name)
    loop0, n, tempco = (int(q) for q in f[f"heat_{tag}_opt"])
    grp = s.temperature_groups() if tempco else None
    x, v = s.x.copy(), s.v.copy()
    ier, _ = Oracle(s).run_cano(x, v, n, 10, 300.0, grp=grp, heat=(loop0, 10, 100.0))
    assert ier == 0
    assert int(f[f"heat_{tag}_loop"]) == loop0 + n
    assert np.array_equal(x, f[f"heat_{tag}_x"]), np.abs(x - f[f"heat_{tag}_x"]).max()
    assert np.array_equal(v, f[f"heat_{tag}_v"])
    x2, v2 = s.x.copy(), s.v.copy()
    assert Oracle(s).run_cano(x2, v2, n, 10, 300.0, grp=grp)[0] == 0
    assert np.abs(x2 - x).max() > 1e-6
