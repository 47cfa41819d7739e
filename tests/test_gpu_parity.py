"""GPU parity through the C ABI (libjpgpu.so) against the oracle and the
golden fixtures of the compiled reference.

Bars: cell assignment and pair lists bit-exact; plain FP64 forces within
1e-12 relative RMS (summation order differs from md_run's); FP32/fixed-point
forces within 1e-5 relative RMS; SHAKE bit-exact against the oracle."""
import numpy as np
import pytest

from conftest import GOLDEN_NAMES, rel_rms
from j_presto_amd import engine, systems
from oracle import Oracle

pytestmark = pytest.mark.gpu


def combined(s, g):
    """grad(:,:,nfrg) as md_run's integrator sees it (verlet.f90:67-91)."""
    if g.shape[0] == 3:
        lam = s.lam
        return g[0] * lam * lam + g[1] * lam + g[2]
    return g[0]


@pytest.mark.parametrize("name", GOLDEN_NAMES)
def test_cells_and_list_bit_exact(golden, name):
    s, f = golden(name)
    print(f"case {name} n={s.n}", flush=True)
    md = engine.MdSystem(s)
    assert md.celprp() == 0
    cell, ptr, idx = md.get_list()
    o = Oracle(s)
    assert o.celprp(s.x) == 0
    ocell, optr, oidx = o.get_list()
    assert np.array_equal(cell, ocell)
    assert np.array_equal(ptr, optr)
    assert np.array_equal(idx, oidx)
    md.close()


@pytest.mark.parametrize("name", GOLDEN_NAMES)
@pytest.mark.parametrize("prec,tol", [(engine.FP64, 1e-12), (engine.FP32, 1e-5)])
def test_forces_energies_vs_reference(golden, name, prec, tol):
    s, f = golden(name)
    print(f"case {name} prec={prec}", flush=True)
    md = engine.MdSystem(s, precision=prec)
    # dnergy: bonds and angles (FP64 in both modes), then cmmdo, into one gradient
    g, et, ier = md.bonded(cord=s.x)
    assert ier == 0
    g, ee, ev, eh, ier = md.cmmdo(grad=g)
    assert ier == 0
    err = rel_rms(combined(s, g), f["force_grad"])
    assert err < tol, err
    etol = 1e-11 if prec == engine.FP64 else 1e-5
    assert abs(ee.sum() - f["force_epot"][8]) <= etol * abs(f["force_epot"][8])
    assert abs(ev.sum() - f["force_epot"][7]) <= etol * abs(f["force_epot"][7])
    for k in range(6):   # bond, angle, torsion, improper, 1-4 vdW, 1-4 elec: FP64 in both modes
        ref = f["force_epot"][1 + k]
        assert abs(et[:, k].sum() - ref) <= 1e-11 * max(abs(ref), 1.0), k
    md.close()


def test_per_class_gradients_lmd(golden):
    """jp_nonbond returns grad(3,N,nfrg) per energy class like cmmdo."""
    s, f = golden("chain_lmd")
    md = engine.MdSystem(s, precision=engine.FP64)
    g, ee, ev, eh, ier = md.cmmdo(cord=s.x)
    o = Oracle(s)
    o.celprp(s.x)
    go, eeo, evo = o.nonbond(s.x)
    for k in range(3):
        assert rel_rms(g[k], go[k]) < 1e-12
        assert abs(ee[k] - eeo[k]) < 1e-11 * abs(eeo).max()
        assert abs(ev[k] - evo[k]) < 1e-11 * abs(evo).max()


@pytest.mark.parametrize("name", ["chain_bonded", "chain_lmd_bonded", "chain_tors", "chain_lmd_tors"])
def test_bonded_per_class_vs_oracle(golden, name):
    """jp_bonded returns grad(3,N,nfrg) and energies per class like
    dnergv_PB's bonded routines, also at displaced coordinates."""
    s, f = golden(name)
    print(f"case bonded {name}", flush=True)
    md = engine.MdSystem(s, precision=engine.FP64)
    o = Oracle(s)
    rng = np.random.default_rng(3)
    for x in (s.x, s.x + rng.normal(scale=0.05, size=s.x.shape)):
        x = np.ascontiguousarray(x)
        g, et, ier = md.bonded(cord=x)
        assert ier == 0
        go, eto = o.bonded(x)
        for k in range(g.shape[0]):
            if np.abs(go[k]).max() == 0.0:
                assert np.abs(g[k]).max() == 0.0
                continue
            assert rel_rms(g[k], go[k]) < 1e-13, k
        assert np.allclose(et, eto, rtol=1e-13, atol=1e-12)
    md.close()


@pytest.mark.parametrize("kind", ["chain", "methyl"])
def test_shake_bit_exact_vs_oracle(kind):
    """jp_shake (2- and 3-atom groups of a chain in water; 4-atom CH3
    groups, shake4) equals the oracle's SHAKE bit for bit."""
    s = systems.chain_in_water(40, 25.0, 5) if kind == "chain" else systems.methyl_water(20, 300, 24.0, 6)
    rng = np.random.default_rng(1)
    x0 = s.x.copy()
    x1 = np.ascontiguousarray(x0 + rng.normal(scale=0.02, size=x0.shape))
    xo = x1.copy()
    o = Oracle(s)
    assert o.shake(xo, x0) == 0
    md = engine.MdSystem(s)
    xg = x1.copy()
    assert md.shake(xg, x0) == 0
    assert np.array_equal(xg, xo)


@pytest.mark.parametrize("name", ["water_1500", "water_6k", "chain_3k", "chain_lmd", "chain_bonded",
                                  "chain_lmd_bonded", "chain_tors", "chain_lmd_tors", "water_42", "methyl_water"])
def test_nve_trajectory_vs_reference(golden, name):
    s, f = golden(name)
    print(f"case {name}", flush=True)
    md = engine.MdSystem(s, precision=engine.FP64)
    md.set_state(s.x, s.v)
    assert md.celprp() == 0
    log, ier = md.runmd(int(f["traj_nsteps"]), update=10)
    assert ier == 0
    x, v = md.get_state()
    assert np.abs(x - f["traj_x"]).max() < 1e-9
    assert np.abs(v - f["traj_v"]).max() < 1e-9


def test_fp32_trajectory_close():
    s, f = load = None, None
    s = systems.water_box(500, 24.8, 7)
    a = engine.MdSystem(s, precision=engine.FP64)
    b = engine.MdSystem(s, precision=engine.FP32)
    la, _ = a.runmd(20, update=10)
    lb, _ = b.runmd(20, update=10)
    xa, va = a.get_state()
    xb, vb = b.get_state()
    assert np.abs(xa - xb).max() < 1e-5


def test_batch_equals_single_runs():
    syss = [systems.water_box(200, 24.2, 21), systems.chain_in_water(48, 24.6, 22, lmd=0.8),
            systems.water_box(700, 28.0, 23, n_ion_pairs=2),
            systems.chain_in_water(64, 24.8, 24, lmd=0.6, bond_fc=300.0, angle_fc=50.0, torsion_fc=1.0)]
    for prec in (engine.FP64, engine.FP32):
        ens = engine.Ensemble(syss, precision=prec)
        log, ier = ens.runmd(20, update=10, nmon=10)
        assert ier == 0
        for i, s in enumerate(syss):
            md = engine.MdSystem(s, precision=prec)
            l1, ier1 = md.runmd(20, update=10, nmon=10)
            x1, v1 = md.get_state()
            xb, vb = ens.get_state(i)
            assert np.array_equal(x1, xb) and np.array_equal(v1, vb), (prec, i)
            assert np.allclose(l1["etot"], log[i]["etot"], rtol=0, atol=1e-9)
            for key in ("ebond", "eangle", "etors", "evdw14", "eelec14"):
                assert np.array_equal(l1[key], log[i][key]), key
            md.close()
        ens.close()


def test_energy_conservation_short():
    """After a velocity-rescaled equilibration, 400 NVE steps keep E_tot
    within a small band in both precisions."""
    s = systems.water_box(500, 24.8, 7)
    spans = {}
    for prec in (engine.FP64, engine.FP32):
        md = engine.MdSystem(s, precision=prec)
        assert engine.equilibrate(md, 600) == 0
        log, ier = md.runmd(400, update=10, nmon=10)
        assert ier == 0
        et = log["etot"]
        spans[prec] = float(et.max() - et.min())
        print(f"prec={prec} etot span {spans[prec]:.4f} kcal/mol, ek {log['ekin'].mean():.1f}", flush=True)
        assert spans[prec] < 2.0, (prec, et)
    assert spans[engine.FP32] < 2.0 * spans[engine.FP64] + 0.2


def test_ier99_excluded_pair_out_of_list():
    s = systems.water_box(100, 24.5, 4)
    x = s.x.copy()
    x[1] += np.array([12.0, 12.0, 0.0])
    md = engine.MdSystem(s)
    assert md.celprp(x) == 99


@pytest.mark.parametrize("prec", [engine.FP32, engine.FP64])
def test_time_pair_leaves_state_and_forces(prec):
    """jp_time_pair: a positive mean duration; positions, velocities and the
    forces of the next evaluation are those without it."""
    s = systems.water_box(500, 24.8, 7)
    md = engine.MdSystem(s, precision=prec)
    md.celprp()
    g0 = md.dnergy()[2]
    x0, v0 = md.get_state()
    t = md.time_pair(5)
    assert 0.0 < t < 1.0
    x1, v1 = md.get_state()
    assert np.array_equal(x0, x1) and np.array_equal(v0, v1)
    assert np.array_equal(md.dnergy()[2], g0)


def test_count_pairs_matches_bruteforce():
    s = systems.water_box(500, 24.8, 7)
    md = engine.MdSystem(s)
    md.celprp()
    d = s.x[:, None, :] - s.x[None, :, :]
    d -= s.box * np.round(d / s.box)
    r2 = (d * d).sum(-1)
    excl = {tuple(p) for p in s.excl14()}
    iu = np.triu_indices(s.n, 1)
    ok = r2[iu] <= 144.0
    n = int(ok.sum()) - sum(1 for (i, j) in excl if r2[i, j] <= 144.0)
    assert md.count_pairs() == n


@pytest.mark.parametrize("prec,bonded", [(engine.FP64, False), (engine.FP32, False), (engine.FP32, True)])
def test_graph_steps_equal_eager(prec, bonded):
    """Plain steps replayed from captured HIP graphs (several replays, with
    rebuilds and energy steps in between) give the same bits as kernels
    launched one by one."""
    s = (systems.chain_in_water(64, 24.8, 24, bond_fc=300.0, angle_fc=50.0, bond_r0=1.53, angle_deg=112.0,
                                torsion_fc=1.0)
         if bonded else systems.water_box(700, 28.0, 23, n_ion_pairs=2))
    print(f"case graphs prec={prec} bonded={bonded}", flush=True)
    out = []
    for graphs in (False, True):
        md = engine.MdSystem(s, precision=prec)
        md.set_graphs(graphs)
        logs = []
        for nst in (25, 13, 40):
            log, ier = md.runmd(nst, update=10, nmon=5)
            assert ier == 0
            logs.append(log["etot"].copy())
        out.append((md.get_state(), logs))
        md.close()
    (x0, v0), l0 = out[0]
    (x1, v1), l1 = out[1]
    assert np.array_equal(x0, x1) and np.array_equal(v0, v1)
    for a, b in zip(l0, l1):
        assert np.array_equal(a, b)


def test_batch_graphs_equal_single_eager():
    syss = [systems.water_box(200, 24.2, 21), systems.water_box(700, 28.0, 23, n_ion_pairs=2)]
    ens = engine.Ensemble(syss, precision=engine.FP32)
    _, ier = ens.runmd(30, update=10, nmon=0)
    assert ier == 0
    for i, s in enumerate(syss):
        md = engine.MdSystem(s, precision=engine.FP32)
        md.set_graphs(False)
        _, ier = md.runmd(30, update=10, nmon=0)
        assert ier == 0
        x1, v1 = md.get_state()
        xb, vb = ens.get_state(i)
        assert np.array_equal(x1, xb) and np.array_equal(v1, vb), i
        md.close()
    ens.close()


def test_launch_shapes_and_autotuner():
    """The list does not depend on the list kernel's wavefronts per block;
    the FP32 forces at 2/4/8 wavefronts per i-cluster agree within the FP32
    bar and each setting is reproducible; the autotuner picks valid shapes."""
    from j_presto_amd import native
    s = systems.water_box(700, 28.0, 23, n_ion_pairs=2)
    print("case launch shapes", flush=True)
    ref = None
    lists = []
    for lw in (1, 2, 4, 8):
        md = engine.MdSystem(s, precision=engine.FP32)
        md.set_option(native.JP_OPT_LIST_WAVES, lw)
        assert md.celprp() == 0
        lists.append(md.get_list())
        md.close()
    for a in lists[1:]:
        assert all(np.array_equal(x, y) for x, y in zip(lists[0], a))
    runs = {}
    for pw in (2, 4, 8):
        out = []
        for rep in range(2):
            md = engine.MdSystem(s, precision=engine.FP32)
            md.set_option(native.JP_OPT_PAIR_WAVES, pw)
            _, ier = md.runmd(20, update=10)
            assert ier == 0
            out.append(md.get_state())
            md.close()
        assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1]), pw
        runs[pw] = out[0]
    for pw in (2, 8):
        assert np.abs(runs[pw][0] - runs[4][0]).max() < 1e-5
    md = engine.MdSystem(s, precision=engine.FP32)
    md.set_option(native.JP_OPT_AUTOTUNE, 1)
    _, ier = md.runmd(20, update=10)
    assert ier == 0
    assert md.get_option(native.JP_OPT_PAIR_WAVES) in (2, 4, 8)
    assert md.get_option(native.JP_OPT_LIST_WAVES) in (1, 2, 4, 8)
    assert np.abs(md.get_state()[0] - runs[4][0]).max() < 1e-5
    md.close()


@pytest.mark.parametrize("prec", [engine.FP64, engine.FP32])
def test_list_overflow_during_overlapped_check(prec):
    """A rebuild whose status is read while the next step segment already
    runs: a segment overflow there restores the state, rebuilds and replays
    the segment; the trajectory is the same bits as without the overflow."""
    from j_presto_amd import native
    s = systems.water_box(700, 28.0, 23, n_ion_pairs=2)
    print(f"case overflow prec={prec}", flush=True)
    out = []
    for force in (False, True):
        md = engine.MdSystem(s, precision=prec)
        _, ier = md.runmd(5, update=10)
        assert ier == 0
        if force:   # the rebuild after step 10 gets segments far too small
            md.set_option(native.JP_OPT_LIST_SEGMENT, 64)
        log, ier = md.runmd(35, update=10, nmon=5)
        assert ier == 0
        out.append((md.get_state(), log["etot"].copy()))
        md.close()
    (x0, v0), e0 = out[0]
    (x1, v1), e1 = out[1]
    assert np.array_equal(x0, x1) and np.array_equal(v0, v1)
    assert np.array_equal(e0, e1)


@pytest.mark.parametrize("name,method", [("water_1500", "micr"), ("chain_lmd", "micr"),
                                         ("water_1500", "cano"), ("chain_lmd", "cano")])
def test_energy_log_vs_oracle_and_reference(golden, name, method):
    """jp_run_nve's monitored energies (caltmp K from (v(t-dt/2)+v(t+dt/2))/2,
    E_pot, E_tot; md.f90:618-628) against the oracle within 1e-12 relative,
    and md_run's printed KINETIC / POTENTIAL / TOTAL (monmd.f90 outtlg, 7
    significant digits) within one unit of the last printed digit.  md_run
    prints the unscaled class sum as POTENTIAL (dnergy.f90:78-79); the log's
    epot/etot carry the lambda-scaled sum under LMD and its per-class terms
    give the unscaled one."""
    from conftest import oracle_monitor
    s, f = golden(name)
    print(f"case energy log {name} {method}", flush=True)
    cano = method == "cano"
    nsteps = int(f["cano_nsteps"] if cano else f["traj_nsteps"])
    grp = None
    if cano and str(f["cano_tempco"]) == "YES":
        grp = s.temperature_groups()
    md = engine.MdSystem(s, precision=engine.FP64)
    md.set_state(s.x, s.v)
    if cano:
        md.set_canonical(300.0, 3, 1e-3, tempco=grp is not None)
    assert md.celprp() == 0
    log, ier = md.runmd(nsteps, update=10, nmon=5)
    assert ier == 0
    md.close()
    o = Oracle(s)
    c = (300.0, grp) if cano else None
    sc = oracle_monitor(o, s, s.x.copy(), s.v.copy(), nsteps, 5, c, scaled=True)
    un = oracle_monitor(Oracle(s), s, s.x.copy(), s.v.copy(), nsteps, 5, c, scaled=False)
    assert np.array_equal(log["step"], sc[:, 0].astype(np.int32))
    for key, col in (("etot", 1), ("epot", 2), ("ekin", 3)):
        assert np.allclose(log[key], sc[:, col], rtol=1e-12, atol=1e-12 * np.abs(sc[:, 2]).max()), key
    terms = ("eelec", "evdw", "ebond", "eangle", "etors", "eimp", "evdw14", "eelec14")
    ep_un = sum(log[k].sum(axis=1) for k in terms)
    assert np.allclose(ep_un, un[:, 2], rtol=1e-12, atol=1e-12 * np.abs(un[:, 2]).max())
    mon = f[f"mon_{method}"]
    mon = mon[mon[:, 0] <= nsteps]
    for val, col in ((log["ekin"], 4), (ep_un, 2), (log["ekin"] + ep_un, 1)):
        ulp = 10.0 ** (np.floor(np.log10(np.abs(mon[:, col]))) - 6)   # one unit of the 7th digit
        assert np.all(np.abs(val - mon[:, col]) <= 1.0001 * ulp), (col, val, mon[:, col])


@pytest.mark.parametrize("prec", [engine.FP64, engine.FP32])
def test_complete_list_option(golden, prec):
    """JP_OPT_CELL_TABLE = 0: on the 42 A box, where md_run's cell table
    drops pairs inside the cutoff, the list is the complete minimum-image set
    (the oracle's brute force) and the forces include the dropped pairs."""
    from j_presto_amd import native
    s, f = golden("water_42")
    print(f"case complete list prec={prec}", flush=True)
    o = Oracle(s)
    bp, bi = o.list_brute(s.x)
    md = engine.MdSystem(s, precision=prec)
    assert md.get_option(native.JP_OPT_CELL_TABLE) == 1
    md.set_option(native.JP_OPT_CELL_TABLE, 0)
    assert md.celprp() == 0
    cell, ptr, idx = md.get_list()
    assert np.array_equal(ptr, bp) and np.array_equal(idx, bi)
    g, ee, ev, eh, ier = md.cmmdo(cord=s.x)
    md.close()
    # the dropped pairs inside the cutoff change the forces by far more than round-off
    assert rel_rms(g[0], f["force_grad"]) > 1e-4


def test_prepared_graphs_are_replayed():
    """jp_prepare_graphs captures every step graph of the following run ahead
    of time: the run captures none and gives the same bits as one without it."""
    from j_presto_amd import native
    s = systems.water_box(700, 28.0, 23, n_ion_pairs=2)
    print("case prepared graphs", flush=True)
    out = []
    for prep in (False, True):
        md = engine.MdSystem(s, precision=engine.FP32)
        _, ier = md.runmd(7, update=10)
        assert ier == 0
        # a rebuild here takes the segment size the first lists asked for, so
        # that no list grows during the run (a grown list needs new graphs)
        assert md.celprp() == 0
        md.set_step(0)
        if prep:
            md.prepare_graphs(60, update=10, nmon=25)
        log, ier = md.runmd(60, update=10, nmon=25)
        assert ier == 0
        assert list(log["step"]) == [25, 50]
        caps = md.get_option(native.JP_OPT_GRAPH_CAPTURES)
        print(f"prepared={prep}: {caps} graphs captured in the run", flush=True)
        assert (caps == 0) if prep else (caps > 0)
        out.append((md.get_state(), log["etot"].copy()))
        md.close()
    assert np.array_equal(out[0][0][0], out[1][0][0]) and np.array_equal(out[0][0][1], out[1][0][1])
    assert np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("name", ["water_1500", "chain_lmd", "chain_lmd_tors", "water_42"])
def test_half_list_option(golden, name):
    """JP_OPT_HALF_LIST = 1: the FP32 path on the half list (each pair once,
    j forces as fixed-point atomics) gives the reference's forces and energies
    within the FP32 bar, exports the same (full) list, and runs batch members
    bit for bit like single runs."""
    from j_presto_amd import native
    s, f = golden(name)
    print(f"case half list {name}", flush=True)
    md = engine.MdSystem(s, precision=engine.FP32)
    md.set_option(native.JP_OPT_HALF_LIST, 1)
    g, et, ier = md.bonded(cord=s.x)
    assert ier == 0
    g, ee, ev, eh, ier = md.cmmdo(cord=s.x, grad=g)
    assert ier == 0
    assert rel_rms(combined(s, g), f["force_grad"]) < 1e-5
    assert abs(ee.sum() - f["force_epot"][8]) <= 1e-5 * abs(f["force_epot"][8])
    assert abs(ev.sum() - f["force_epot"][7]) <= 1e-5 * abs(f["force_epot"][7])
    cell, ptr, idx = md.get_list()
    o = Oracle(s)
    assert o.celprp(s.x) == 0
    ocell, optr, oidx = o.get_list()
    assert np.array_equal(ptr, optr) and np.array_equal(idx, oidx)
    md.close()
    # trajectories: single and batched runs equal bit for bit, close to the full list's
    runs = []
    for half in (0, 1):
        m2 = engine.MdSystem(s, precision=engine.FP32)
        m2.set_option(native.JP_OPT_HALF_LIST, half)
        log, ier = m2.runmd(20, update=10, nmon=10)
        assert ier == 0
        runs.append((m2.get_state(), log["etot"].copy()))
        m2.close()
    assert np.abs(runs[0][0][0] - runs[1][0][0]).max() < 1e-5
    ens = engine.Ensemble([systems.water_box(200, 24.2, 21), s], precision=engine.FP32)
    ens.set_option(native.JP_OPT_HALF_LIST, 1)
    log, ier = ens.runmd(20, update=10, nmon=10)
    assert ier == 0
    xb, vb = ens.get_state(1)
    ens.close()
    assert np.array_equal(xb, runs[1][0][0]) and np.array_equal(vb, runs[1][0][1])
    assert np.array_equal(log[1]["etot"], runs[1][1])
