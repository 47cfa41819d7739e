"""METHOD= CANO on the GPU (jp_set_canonical) against md_run's canonical
trajectories (golden fixtures) and the oracle.

Bars: FP64 trajectories within 1e-9 (A, A/fs) of md_run's after 10-20 steps
(the group kinetic-energy sums run in a fixed tree order, md_run's serially,
so the last bits differ); runs deterministic, graph replays equal to host
launches and batch members equal to single runs bit for bit."""
import numpy as np
import pytest

from conftest import ge_params
from j_presto_amd import engine, native, systems
from oracle import Oracle

pytestmark = pytest.mark.gpu

CANO_NAMES = ["water_1500", "chain_3k", "chain_lmd", "methyl_water"]


@pytest.mark.parametrize("name", CANO_NAMES)
def test_cano_trajectory_vs_reference(golden, name):
    s, f = golden(name)
    print(f"case {name}", flush=True)
    md = engine.MdSystem(s, precision=engine.FP64)
    md.set_state(s.x, s.v)
    md.set_canonical(300.0, tempco=str(f["cano_tempco"]) == "YES")
    assert md.celprp() == 0
    _, ier = md.runmd(int(f["cano_nsteps"]), update=10)
    assert ier == 0
    x, v = md.get_state()
    assert np.abs(x - f["cano_x"]).max() < 1e-9
    assert np.abs(v - f["cano_v"]).max() < 1e-9
    md.close()


@pytest.mark.parametrize("tempco", [False, True])
def test_cano_vs_oracle_fp32_close(tempco):
    """FP32/fixed-point pair forces in the canonical step: within 1e-5 A of
    the oracle's FP64 canonical trajectory after 20 steps."""
    s = systems.chain_in_water(48, 24.6, 31)
    grp = s.temperature_groups() if tempco else None
    o = Oracle(s)
    x, v = s.x.copy(), s.v.copy()
    assert o.celprp(x) == 0
    for idloop in range(1, 21):
        g, _ = o.bonded(x)
        g, _, _ = o.nonbond(x, g)
        assert o.cano_step(x, v, o.combine(g), 300.0, 3, 1e-3, grp)[0] == 0
        if idloop % 10 == 0:
            assert o.celprp(x) == 0
    md = engine.MdSystem(s, precision=engine.FP32)
    md.set_canonical(300.0, tempco=tempco)
    _, ier = md.runmd(20, update=10)
    assert ier == 0
    xg, vg = md.get_state()
    assert np.abs(xg - x).max() < 1e-5
    md.close()


def test_cano_graphs_equal_eager_and_batch():
    """Graph replays equal host launches, and batch members equal single
    runs, bit for bit (energy steps included)."""
    syss = [systems.water_box(200, 24.2, 41), systems.chain_in_water(48, 24.6, 42, lmd=0.8)]
    ens = engine.Ensemble(syss, precision=engine.FP32)
    ens.set_canonical(310.0, tempco=True)
    log, ier = ens.runmd(25, update=10, nmon=5)
    assert ier == 0
    for i, s in enumerate(syss):
        res = []
        for graphs in (True, False):
            md = engine.MdSystem(s, precision=engine.FP32)
            md.set_graphs(graphs)
            md.set_canonical(310.0, tempco=True)
            l1, ier1 = md.runmd(25, update=10, nmon=5)
            assert ier1 == 0
            res.append(md.get_state() + (l1["etot"],))
            md.close()
        (xa, va, ea), (xb, vb, eb) = res
        assert np.array_equal(xa, xb) and np.array_equal(va, vb)
        assert np.array_equal(ea, eb)
        xe, ve = ens.get_state(i)
        assert np.array_equal(xa, xe) and np.array_equal(va, ve), i
        assert np.array_equal(ea, log[i]["etot"])
    ens.close()


def test_cano_reaches_target_temperature():
    """Water started at 300 K and coupled to 360 K: after 100 steps the
    temperature of 1/2 (v- + v+) is within 2 % of the target, while the same
    run in METHOD= MICR stays near its start."""
    s = systems.water_box(500, 24.8, 7)
    dof = Oracle(s).group_dof()[0]
    temps = {}
    for cano in (True, False):
        md = engine.MdSystem(s, precision=engine.FP64)
        if cano:
            md.set_canonical(360.0)
        log, ier = md.runmd(100, update=10, nmon=100)
        assert ier == 0
        # caltmp.f90: T = 2 Ek / (dof R) with Ek in kcal/mol
        temps[cano] = 2.0 * log["ekin"][-1] * 4184.0 / (dof * 8.31451)
        md.close()
    assert abs(temps[True] - 360.0) < 0.02 * 360.0
    assert abs(temps[False] - 360.0) > 0.05 * 360.0


def test_cano_off_returns_to_micr():
    s = systems.water_box(200, 24.2, 43)
    a = engine.MdSystem(s, precision=engine.FP64)
    a.set_canonical(300.0)
    a.set_canonical(None)
    b = engine.MdSystem(s, precision=engine.FP64)
    for md in (a, b):
        _, ier = md.runmd(12, update=10)
        assert ier == 0
    assert all(np.array_equal(p, q) for p, q in zip(a.get_state(), b.get_state()))
    with pytest.raises(Exception):
        a.set_canonical(-1.0)
    a.close()
    b.close()


def _no_shake(s):
    for m in s.mols:
        m.shake = []
    return s


def test_cano_without_shake_groups():
    """A system without SHAKE groups takes chi = 1 (verlethe.f90:396-404)
    and no rounds: FP64 against the oracle within 1e-9 after 10 steps, alone
    (begin + drift kernels) and in a batch with a SHAKE system (the drift
    inside the first round's kernel), the batch member equal to the single
    run bit for bit."""
    s = _no_shake(systems.water_box(200, 24.2, 44))
    assert not s.shake_groups()
    o = Oracle(s)
    x, v = s.x.copy(), s.v.copy()
    assert o.celprp(x) == 0
    for idloop in range(1, 11):
        g, _ = o.bonded(x)
        g, _, _ = o.nonbond(x, g)
        assert o.cano_step(x, v, o.combine(g), 300.0, 3, 1e-3, None)[0] == 0
        if idloop % 10 == 0:
            assert o.celprp(x) == 0
    md = engine.MdSystem(s, precision=engine.FP64)
    md.set_canonical(300.0)
    _, ier = md.runmd(10, update=10)
    assert ier == 0
    xs, vs = md.get_state()
    md.close()
    assert np.abs(xs - x).max() < 1e-9 and np.abs(vs - v).max() < 1e-9
    ens = engine.Ensemble([systems.water_box(200, 24.2, 45), s], precision=engine.FP64)
    ens.set_canonical(300.0)
    _, ier = ens.runmd(10, update=10)
    assert ier == 0
    xe, ve = ens.get_state(1)
    ens.close()
    assert np.array_equal(xe, xs) and np.array_equal(ve, vs)


@pytest.mark.parametrize("graphs", [True, False])
def test_cano_shake_retry_vs_reference(golden, graphs):
    """METHOD= CANO with LIMSHK= 3: SHAKE fails at the full step, and
    verlethe undoes the substep, halves dt and redoes it (verlethe.f90:
    552-573, 745-752).  md_run's restart set-up step (inirst) runs first: it
    leaves the velocities of its last substep's start (md.f90:1240-1250).
    The trajectory matches md_run's within 1e-9, with step graphs (a failed
    segment is redone step by step) and without."""
    s, f = golden("water_1500")
    print(f"case cano retry graphs={graphs}", flush=True)
    assert int(f["retry_provisions"]) > 0
    md = engine.MdSystem(s, precision=engine.FP64, shake_maxit=3)
    md.set_graphs(graphs)
    md.set_state(s.x, s.v)
    md.set_canonical(300.0, 3, 1e-3)
    assert md.celprp() == 0
    rlog, ier = md.restart_step()
    assert ier == 0 and rlog["step"][0] == 0
    x0, _ = md.get_state()
    assert np.array_equal(x0, s.x)
    log, ier = md.runmd(int(f["retry_loop"]), update=10)
    assert ier == 0
    assert md.get_step() == int(f["retry_loop"])
    x, v = md.get_state()
    md.close()
    print(f"max |dx| {np.abs(x - f['retry_x']).max():.2e}, max |dv| {np.abs(v - f['retry_v']).max():.2e}",
          flush=True)
    assert np.abs(x - f["retry_x"]).max() < 1e-9
    assert np.abs(v - f["retry_v"]).max() < 1e-9


def test_restart_step_keeps_the_state_without_a_retry(golden):
    """Without a SHAKE failure the restart set-up step leaves coordinates,
    velocities and the step counter as they were, and its record is the
    energy of the restart state (E_pot of the oracle's forces at x)."""
    s, f = golden("water_1500")
    print("case restart step", flush=True)
    md = engine.MdSystem(s, precision=engine.FP64)
    md.set_state(s.x, s.v)
    md.set_step(25)
    assert md.celprp() == 0
    rlog, ier = md.restart_step()
    assert ier == 0 and rlog["step"][0] == 25 and md.get_step() == 25
    x, v = md.get_state()
    md.close()
    assert np.array_equal(x, s.x) and np.array_equal(v, s.v)
    o = Oracle(s)
    o.celprp(s.x)
    g, et = o.bonded(s.x)
    g, ee, ev = o.nonbond(s.x, g)
    epot = et.sum() + ee.sum() + ev.sum()
    assert abs(rlog["epot"][0] - epot) < 1e-9 * max(1.0, abs(epot))


@pytest.mark.parametrize("tag,graphs", [("win", True), ("win", False), ("lo", True), ("hi", True)])
def test_ge_vs_reference(golden, tag, graphs):
    """METHOD= GE (multicanonical, verlethe.f90:196-233), FP64: every step's
    forces carry the weight of its potential energy, formed on the device
    (window polynomials for "win", whose energy crosses all three windows;
    alphalw for "lo", alphaup for "hi").  After md_run's restart step the
    trajectory matches md_run's within 1e-9, with step graphs and without."""
    s, f = golden("water_1500")
    print(f"case ge {tag} graphs={graphs}", flush=True)
    md = engine.MdSystem(s, precision=engine.FP64)
    md.set_graphs(graphs)
    md.set_state(s.x, s.v)
    md.set_canonical(300.0, 3, 1e-3)
    md.set_mcmd(ge_params(f, tag))
    assert md.celprp() == 0
    _, ier = md.restart_step()
    assert ier == 0
    _, ier = md.runmd(int(f[f"ge_{tag}_loop"]), update=10)
    assert ier == 0
    x, v = md.get_state()
    md.close()
    print(f"max |dx| {np.abs(x - f[f'ge_{tag}_x']).max():.2e}", flush=True)
    assert np.abs(x - f[f"ge_{tag}_x"]).max() < 1e-9
    assert np.abs(v - f[f"ge_{tag}_v"]).max() < 1e-9


@pytest.mark.parametrize("prec,tol", [(engine.FP64, 1e-9), (engine.FP32, 1e-5)])
def test_ge_with_bonded_terms_vs_oracle(prec, tol):
    """METHOD= GE on a chain in water: the device energy includes the bonded
    terms per slot.  FP64 within 1e-9 A of Oracle.run_cano with the same
    parameters after 20 steps (energy window around the start), FP32 within
    1e-5 A."""
    s = systems.chain_in_water(48, 24.6, 31)
    o = Oracle(s)
    assert o.celprp(s.x) == 0
    g, et = o.bonded(s.x)
    g, ee, ev = o.nonbond(s.x, g)
    e0 = et[0].sum() + ev[0] + ee[0]
    ge = dict(c=np.array([[1.70, -0.3, 0.2], [1.65, 0.1, -0.1]]), low=np.array([0.0, 0.5]),
              high=np.array([0.5, 1.0]), alphalw=1.66, alphaup=1.69, celw=e0 - 3000.0, ceup=e0 + 1000.0, tempce=300.0)
    x, v = s.x.copy(), s.v.copy()
    o = Oracle(s)
    assert o.run_cano(x, v, 20, 10, 300.0, 3, 1e-3, ge=ge) == (0, 0)
    print(f"case ge chain prec={prec}", flush=True)
    md = engine.MdSystem(s, precision=prec)
    md.set_canonical(300.0, 3, 1e-3)
    md.set_mcmd(ge)
    assert md.celprp() == 0
    assert md.restart_step()[1] == 0
    _, ier = md.runmd(20, update=10)
    assert ier == 0
    xg, vg = md.get_state()
    md.close()
    print(f"max |dx| {np.abs(xg - x).max():.2e}", flush=True)
    assert np.abs(xg - x).max() < tol


def test_ge_requires_one_class_and_canonical(golden):
    """METHOD= GE takes one system with one energy class and runs only with
    the canonical temperature control on."""
    s, f = golden("water_1500")
    ge = ge_params(f, "win")
    md = engine.MdSystem(s, precision=engine.FP64)
    md.set_mcmd(ge)
    with pytest.raises(native.JpError):
        md.runmd(2, update=10)
    md.set_canonical(300.0, 3, 1e-3)
    assert md.runmd(2, update=10)[1] == 0
    md.set_mcmd(None)
    md.close()
    lmd = systems.chain_in_water(48, 24.6, 31, lmd=0.8)
    md = engine.MdSystem(lmd, precision=engine.FP64)
    with pytest.raises(native.JpError):
        md.set_mcmd(ge)
    md.close()


@pytest.mark.parametrize("name", ["water_1500", "chain_lmd"])
@pytest.mark.parametrize("tag", ["a", "b"])
def test_heating_loop_vs_reference(golden, name, tag):
    """md_run's heating loop (HEATLO= 10, STARTT= 100 K to SETTEM= 300 K)
    from restarts at loop 0 and 4: the ramp's steps as host launches, the
    rest as graphs; x and v within 1e-9 of md_run's."""
    s, f = golden(name)
    loop0, n, tempco = (int(q) for q in f[f"heat_{tag}_opt"])
    print(f"case heating {name} {tag}", flush=True)
    md = engine.MdSystem(s, precision=engine.FP64)
    md.set_state(s.x, s.v)
    md.set_canonical(300.0, tempco=bool(tempco))
    md.set_heating(10, 100.0)
    md.set_step(loop0)
    assert md.celprp() == 0
    _, ier = md.restart_step()
    assert ier == 0
    _, ier = md.runmd(n, update=10)
    assert ier == 0
    assert md.get_step() == loop0 + n
    x, v = md.get_state()
    md.close()
    dx, dv = np.abs(x - f[f"heat_{tag}_x"]).max(), np.abs(v - f[f"heat_{tag}_v"]).max()
    print(f"max |dx| {dx:.2e} |dv| {dv:.2e}", flush=True)
    assert dx < 1e-9 and dv < 1e-9
