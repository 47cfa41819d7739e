"""The cell multipole method (md_run CAL15M= CMM, BOUNDA= NO) on the GPU
through the C ABI (jp_set_cmm), against md_run's force dumps and MICR runs
(fixtures droplet_cmm, chain_droplet_cmm, chain_droplet_lmd) and the oracle.

Bars: level-1 cells equal to those of the coordinate mean; FP64 forces
within 1e-12 relative RMS and energies within 1e-11 relative of md_run
(summation order differs); MICR trajectories with the far field every
CMMUPD= 5 steps within 1e-9 of md_run's; step graphs equal host launches
bitwise; md_run's ier (28, 51, 99) equal to the oracle's."""
import numpy as np
import pytest

from conftest import rel_rms
from j_presto_amd import engine
from oracle import Oracle

pytestmark = pytest.mark.gpu

CMM_NAMES = ["droplet_cmm", "chain_droplet_cmm", "chain_droplet_lmd"]


def combined(s, g):
    if g.shape[0] == 3:
        lam = s.lam
        return g[0] * lam * lam + g[1] * lam + g[2]
    return g[0]


def cmm_md(s, graphs=True, cminsiz=6.0, nlev=3):
    md = engine.MdSystem(s, precision=engine.FP64)
    md.set_graphs(graphs)
    md.set_state(s.x, s.v)
    md.set_cmm(cminsiz, nlev, 5)
    return md


@pytest.mark.parametrize("name", CMM_NAMES)
def test_cmm_cells(golden, name):
    """celpar_CB: cells of 6 A centred on the coordinate mean, 16 per edge."""
    s, _ = golden(name)
    print(f"case cmm cells {name} n={s.n}", flush=True)
    md = cmm_md(s)
    assert md.celprp() == 0
    cell = md.cmm_cells()
    md.close()
    nv, siz = 16, 6.0
    c0 = s.x.sum(0) * (1.0 / s.n)
    org = (siz * (1 - 0.5 * (nv + 1)) + c0) - 0.5 * siz
    f = (s.x - org) / siz
    assert np.array_equal(cell, f.astype(np.int32) + 1)


@pytest.mark.parametrize("name", CMM_NAMES)
def test_cmm_forces_vs_reference(golden, name):
    s, f = golden(name)
    print(f"case cmm forces {name}", flush=True)
    md = cmm_md(s)
    assert md.celprp() == 0
    g, et, ier = md.bonded(cord=s.x)
    assert ier == 0
    g, ee, ev, eh, ier = md.cmmdo(grad=g)
    assert ier == 0
    md.close()
    err = rel_rms(combined(s, g), f["force_grad"])
    print(f"rel rms {err:.2e}  elec {ee.sum()!r} vs {f['force_epot'][8]!r}", flush=True)
    assert err < 1e-12, err
    assert abs(ee.sum() - f["force_epot"][8]) <= 1e-11 * abs(f["force_epot"][8])
    assert abs(ev.sum() - f["force_epot"][7]) <= 1e-11 * abs(f["force_epot"][7])
    assert not eh.any()


@pytest.mark.parametrize("name", CMM_NAMES)
def test_cmm_classes_vs_oracle(golden, name):
    """Per-class gradients and energies at moved coordinates (a fresh far
    field) against the oracle's CMM."""
    s, _ = golden(name)
    print(f"case cmm classes {name}", flush=True)
    x = s.x + np.random.default_rng(11).normal(scale=0.03, size=s.x.shape)
    md = cmm_md(s)
    assert md.celprp(cord=x) == 0
    g, ee, ev, _, ier = md.cmmdo()
    assert ier == 0
    md.close()
    o = Oracle(s)
    assert o.cmm_setup(s.x) == 0 and o.cmm_celprp(x) == 0
    go, eeo, evo = o.cmm_nonbond(x)
    for k in range(g.shape[0]):
        assert rel_rms(g[k], go[k]) < 1e-12, k
    assert np.allclose(ee, eeo, rtol=1e-11, atol=0) and np.allclose(ev, evo, rtol=1e-11, atol=0)


@pytest.mark.parametrize("name", CMM_NAMES)
def test_cmm_micr_trajectory_vs_reference(golden, name):
    s, f = golden(name)
    n = int(f["traj_nsteps"])
    print(f"case cmm micr {name} {n} steps", flush=True)
    md = cmm_md(s)
    assert md.celprp() == 0
    log, ier = md.runmd(n, update=10)
    assert ier == 0
    x, v = md.get_state()
    md.close()
    dx, dv = np.abs(x - f["traj_x"]).max(), np.abs(v - f["traj_v"]).max()
    print(f"max |dx| {dx:.2e} |dv| {dv:.2e}", flush=True)
    assert dx < 1e-9 and dv < 1e-9


def test_cmm_graphs_equal_eager(golden):
    """Graph segments replay the host-launched steps bit for bit, the far
    field's reuse pattern included (25 steps: segments start at every phase
    of CMMUPD)."""
    s, _ = golden("chain_droplet_lmd")
    print("case cmm graphs/eager", flush=True)
    out = []
    for graphs in (True, False):
        md = cmm_md(s, graphs=graphs)
        assert md.celprp() == 0
        _, ier = md.runmd(7, update=10)
        assert ier == 0
        _, ier = md.runmd(18, update=10, nmon=4)
        assert ier == 0
        out.append(md.get_state())
        md.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("cminsiz", [1.0, 1.5, 2.0, 3.0])
def test_cmm_ier_matches_oracle(golden, cminsiz):
    """Small CELSIZ: md_run's ier of the rebuild (28 too many atoms in a
    cell, 99 excluded pairs beyond the tables, or 0) as the oracle's."""
    s, _ = golden("droplet_cmm")
    print(f"case cmm ier cminsiz={cminsiz}", flush=True)
    o = Oracle(s)
    assert o.cmm_setup(s.x, cminsiz, 3) == 0
    ref = o.cmm_celprp(s.x)
    md = cmm_md(s, cminsiz=cminsiz)
    ier = md.celprp()
    md.close()
    print(f"ier {ier} oracle {ref}", flush=True)
    assert ier == ref


@pytest.mark.parametrize("name", ["droplet_cmm", "chain_droplet_lmd"])
def test_cmm_cano_vs_reference(golden, name):
    """METHOD= CANO (TEMPCO= NO) under CMM: 12 steps across a rebuild and two
    far-field recomputations."""
    s, f = golden(name)
    print(f"case cmm cano {name}", flush=True)
    md = cmm_md(s)
    md.set_canonical(300.0)
    assert md.celprp() == 0
    _, ier = md.runmd(int(f["cano_nsteps"]), update=10)
    assert ier == 0
    x, v = md.get_state()
    md.close()
    dx, dv = np.abs(x - f["cano_x"]).max(), np.abs(v - f["cano_v"]).max()
    print(f"max |dx| {dx:.2e} |dv| {dv:.2e}", flush=True)
    assert dx < 1e-9 and dv < 1e-9


def test_cmm_alsd_vs_reference(golden):
    """ALSD (METHOD= GE under GEPSME= LMD) under CMM: lambda's force from the
    near and far class energies of every step."""
    from conftest import alsd_params
    s, f = golden("chain_droplet_lmd")
    print("case cmm alsd", flush=True)
    al, tempco = alsd_params(s, f, "a")
    md = cmm_md(s)
    md.set_canonical(300.0, 3, 1e-3, tempco=tempco)
    md.set_alsd(al)
    md.set_lambda(s.lam, 0.0)
    assert md.celprp() == 0
    _, ier = md.restart_step()
    assert ier == 0
    _, ier = md.runmd(int(f["alsd_a_loop"]), update=10)
    assert ier == 0
    x, v = md.get_state()
    lam, lamv = md.get_lambda()
    md.close()
    print(f"max |dx| {np.abs(x - f['alsd_a_x']).max():.2e} dlam {lam - float(f['alsd_a_lam']):.2e}", flush=True)
    assert np.abs(x - f["alsd_a_x"]).max() < 1e-9
    assert np.abs(v - f["alsd_a_v"]).max() < 1e-9
    assert abs(lam - float(f["alsd_a_lam"])) < 1e-11
    assert abs(lamv - float(f["alsd_a_lamv"])) < 1e-11


def test_cmm_energy_log_vs_reference(golden):
    """200 MICR steps of the droplet under CMM from its generated (strained)
    state, which heats to about 1000 K: the monitored E_tot, E_pot and K
    every 20 steps against md_run's printed ones (7 significant digits)
    within one unit of the last digit.  md_run's E_tot itself rises by
    about 370 kcal/mol over these steps (far field and tables held between
    updates); the GPU run follows it."""
    s, f = golden("droplet_cmm")
    print("case cmm energy log", flush=True)
    md = cmm_md(s)
    assert md.celprp() == 0
    log, ier = md.runmd(200, update=10, nmon=20)
    assert ier == 0
    md.close()
    mon = f["mon_micr"]
    mon = mon[(mon[:, 0] <= 200) & (mon[:, 0] % 20 == 0)]
    assert np.array_equal(log["step"], mon[:, 0].astype(np.int32))
    for key, col in (("etot", 1), ("epot", 2), ("ekin", 4)):
        ulp = 10.0 ** (np.floor(np.log10(np.abs(mon[:, col]))) - 6)
        d = np.abs(log[key] - mon[:, col])
        print(key, (d / ulp).max(), flush=True)
        assert (d <= 1.0001 * ulp).all(), key
