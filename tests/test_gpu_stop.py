"""Centre-of-mass stopping (jp_set_stopce, md_run STOPCE) and md_run's initial
step of a run without a restart (jp_start_step, staver) on the GPU against
md_run's runs (STOP_RUNS fixtures of tests/golden/make_golden.py
--stop-only) and the oracle.

Bars: FP64 x and v within 1e-9 (A, A/fs) of md_run's after 12 steps; the
block sums of remotv and of the canonical temperatures run in fixed tree
orders, md_run's serially, so the last bits differ.  STOPCE= TRAN, whose
md_run run reads an unset angular velocity, against the oracle; graph
replays equal to host launches and batch members equal to single runs bit
for bit."""
import numpy as np
import pytest

from j_presto_amd import engine
from oracle import Oracle

pytestmark = pytest.mark.gpu

STOP_NAMES = ["chain_tors", "chain_lmd_tors", "chain_droplet_cmm"]
# tag -> (METHOD, STOPCE, RESTAR, steps, TEMPCO), as the fixtures were made
STOP_RUNS = {"mb": ("MICR", "BOTH", "YES", 12, "NO"), "mr": ("MICR", "ROTA", "YES", 12, "NO"),
             "cb": ("CANO", "BOTH", "YES", 12, "NO"), "cy": ("CANO", "ROTA", "YES", 12, "YES"),
             "s0": ("MICR", "NO", "NO", 12, "NO"), "sc": ("CANO", "BOTH", "NO", 12, "NO")}


def stop_md(s, cmm, method, stopce, restar, tempco, graphs=True, precision=engine.FP64):
    md = engine.MdSystem(s, precision=precision)
    md.set_graphs(graphs)
    v = s.v if restar == "YES" else np.zeros_like(s.v)   # INITIA= ZERO
    md.set_state(s.x, v)
    if cmm:
        md.set_cmm(6.0, 3, 5)
    md.set_stopce(stopce)
    if method == "CANO":
        md.set_canonical(300.0, tempco=tempco == "YES")
    return md


def run_stop(s, tag, cmm, graphs=True):
    method, stopce, restar, n, tempco = STOP_RUNS[tag]
    md = stop_md(s, cmm, method, stopce, restar, tempco, graphs)
    assert md.celprp() == 0
    if restar == "YES":
        if method == "CANO":
            _, ier = md.restart_step()
            assert ier == 0
    else:
        assert md.start_step() == 0
        assert md.get_step() == 0
    _, ier = md.runmd(n, update=10)
    assert ier == 0
    x, v = md.get_state()
    md.close()
    return x, v


@pytest.mark.parametrize("name", STOP_NAMES)
@pytest.mark.parametrize("tag", list(STOP_RUNS))
def test_stop_runs_vs_reference(golden, name, tag):
    s, f = golden(name)
    cmm = name == "chain_droplet_cmm"
    print(f"case stop {name} {tag} {STOP_RUNS[tag]}", flush=True)
    x, v = run_stop(s, tag, cmm)
    dx, dv = np.abs(x - f[f"stop_{tag}_x"]).max(), np.abs(v - f[f"stop_{tag}_v"]).max()
    print(f"max |dx| {dx:.2e} |dv| {dv:.2e}", flush=True)
    assert dx < 1e-9 and dv < 1e-9


@pytest.mark.parametrize("method", ["MICR", "CANO"])
def test_stop_tran_vs_oracle(golden, method):
    """STOPCE= TRAN against the oracle (md_run's own TRAN runs read an unset
    angular velocity)."""
    s, _ = golden("chain_tors")
    print(f"case stop TRAN {method}", flush=True)
    md = stop_md(s, False, method, "TRAN", "YES", "NO")
    assert md.celprp() == 0
    if method == "CANO":
        assert md.restart_step()[1] == 0
    assert md.runmd(12, update=10)[1] == 0
    x, v = md.get_state()
    md.close()
    o = Oracle(s)
    assert o.set_stopce("TRAN", s.chain_end(1), s.first_kind_atoms()) == 0
    xo, vo = s.x.copy(), s.v.copy()
    if method == "MICR":
        assert o.run_nve(xo, vo, 12, 10) == 0
    else:
        assert o.run_cano(xo, vo, 12, 10, 300.0)[0] == 0
    dx, dv = np.abs(x - xo).max(), np.abs(v - vo).max()
    print(f"max |dx| {dx:.2e} |dv| {dv:.2e}", flush=True)
    assert dx < 1e-9 and dv < 1e-9
    m = s.mass[:s.chain_end(1), None]
    p = (m * v[:s.chain_end(1)]).sum(0)
    print("chain momentum", np.abs(p).max(), flush=True)


@pytest.mark.parametrize("tag", ["mb", "cb", "sc"])
def test_stop_graphs_equal_eager(golden, tag):
    """Step graphs with the stop kernels replay the host-launched steps bit
    for bit."""
    s, _ = golden("chain_tors")
    print(f"case stop graphs/eager {tag}", flush=True)
    a = run_stop(s, tag, False, graphs=True)
    b = run_stop(s, tag, False, graphs=False)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_stop_batch_equals_single(golden):
    """A batch of two systems, each stopping its own first chain, gives each
    system's single run bit for bit (STOPCE= BOTH under METHOD= CANO, from a
    start without a restart)."""
    from j_presto_amd import systems
    syss = [golden("chain_tors")[0], systems.chain_in_water(48, 24.6, 42)]
    print("case stop batch", flush=True)
    ens = engine.Ensemble(syss, precision=engine.FP64)
    for i, s in enumerate(syss):
        ens.set_state(i, s.x, np.zeros_like(s.v))
    ens.set_stopce("BOTH")
    ens.set_canonical(300.0)
    assert ens.start_step() == 0
    assert ens.runmd(8, update=10)[1] == 0
    for i, s in enumerate(syss):
        md = stop_md(s, False, "CANO", "BOTH", "NO", "NO")
        assert md.celprp() == 0 and md.start_step() == 0
        assert md.runmd(8, update=10)[1] == 0
        x, v = md.get_state()
        md.close()
        xe, ve = ens.get_state(i)
        assert np.array_equal(x, xe) and np.array_equal(v, ve), i
    ens.close()


def test_stop_errors_and_refusals(golden):
    """A stopped first atom alone, exactly at the origin (the system shifted
    by its coordinates), has no inertia: the step ends with
    JP_IER_STOP_SING (-3; md_run's remotv gives -2).  set_stopce after
    set_canonical, an unknown mode or a chain end beyond the system are
    refused."""
    from j_presto_amd import native
    s, _ = golden("chain_tors")
    print("case stop errors", flush=True)
    md = engine.MdSystem(s, precision=engine.FP64)
    md.set_state(s.x - s.x[0], s.v)
    assert md.lib.jp_set_stopce(md.ctx, 3, 1, 0) == 0
    assert md.celprp() == 0
    assert md.runmd(1, update=10)[1] == -3
    md.close()
    md = engine.MdSystem(s, precision=engine.FP64)
    assert md.lib.jp_set_stopce(md.ctx, 4, 10, 0) == native.JP_EINVAL
    assert md.lib.jp_set_stopce(md.ctx, 1, s.n + 1, 0) == native.JP_EINVAL
    md.set_canonical(300.0)
    assert md.lib.jp_set_stopce(md.ctx, 1, 10, 0) == native.JP_EINVAL
    md.close()
