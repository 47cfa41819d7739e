"""The autotuner's knobs through the C ABI: box-grid cell size
(JP_OPT_GRID_SHIFT), the list kernel's tile width (JP_OPT_LIST_BATCH) and the
tuner's choice of them.  None of them may change the pair list (bit-exact
against the oracle, which restates cutatm_PB, cutoff.f90:166-505); the grid
regroups partial sums, so forces stay within the bars of
test_gpu_parity.py (FP64 1e-12, FP32/fixed point 1e-5 relative RMS against
md_run's vectors)."""
import numpy as np
import pytest

from conftest import rel_rms
from j_presto_amd import engine, native, systems
from oracle import Oracle

pytestmark = pytest.mark.gpu


def combined(s, g):
    if g.shape[0] == 3:
        lam = s.lam
        return g[0] * lam * lam + g[1] * lam + g[2]
    return g[0]


@pytest.mark.parametrize("name", ["water_6k", "water_54", "chain_lmd_tors", "water_42"])
@pytest.mark.parametrize("shift,nb", [(-1, 2), (1, 2), (0, 1), (0, 4), (1, 4)])
def test_grid_and_tile_width_keep_list_and_forces(golden, name, shift, nb):
    s, f = golden(name)
    print(f"case {name} grid shift {shift} list batches {nb}", flush=True)
    o = Oracle(s)
    assert o.celprp(s.x) == 0
    ocell, optr, oidx = o.get_list()
    for prec, tol in ((engine.FP64, 1e-12), (engine.FP32, 1e-5)):
        md = engine.MdSystem(s, precision=prec)
        md.set_option(native.JP_OPT_GRID_SHIFT, shift)
        md.set_option(native.JP_OPT_LIST_BATCH, nb)
        assert md.get_option(native.JP_OPT_GRID_SHIFT) == shift
        assert md.get_option(native.JP_OPT_LIST_BATCH) == nb
        assert md.celprp() == 0
        cell, ptr, idx = md.get_list()
        assert np.array_equal(cell, ocell)
        assert np.array_equal(ptr, optr)
        assert np.array_equal(idx, oidx)
        g, et, ier = md.bonded(cord=s.x)
        assert ier == 0
        g, ee, ev, eh, ier = md.cmmdo(grad=g)
        assert ier == 0
        err = rel_rms(combined(s, g), f["force_grad"])
        assert err < tol, (prec, err)
        etol = 1e-11 if prec == engine.FP64 else 1e-5
        assert abs(ee.sum() - f["force_epot"][8]) <= etol * abs(f["force_epot"][8])
        assert abs(ev.sum() - f["force_epot"][7]) <= etol * abs(f["force_epot"][7])
        md.close()


def test_grid_shift_run_is_reproducible_and_close():
    """A run on a finer grid is bitwise reproducible and stays within the
    FP32 bar of the default grid's run; switching the grid back restores the
    default run's bits."""
    s = systems.water_box(700, 28.0, 23, n_ion_pairs=2)
    print("case grid-shift runs", flush=True)

    def run(shift):
        md = engine.MdSystem(s, precision=engine.FP32)
        if shift is not None:
            md.set_option(native.JP_OPT_GRID_SHIFT, shift)
        _, ier = md.runmd(20, update=10)
        assert ier == 0
        out = md.get_state()
        md.close()
        return out

    base = run(None)
    a, b = run(1), run(1)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert np.abs(a[0] - base[0]).max() < 1e-5
    md = engine.MdSystem(s, precision=engine.FP32)
    md.set_option(native.JP_OPT_GRID_SHIFT, 1)
    md.set_option(native.JP_OPT_GRID_SHIFT, 0)
    _, ier = md.runmd(20, update=10)
    assert ier == 0
    x, v = md.get_state()
    md.close()
    assert np.array_equal(x, base[0]) and np.array_equal(v, base[1])


@pytest.mark.parametrize("prec", [engine.FP64, engine.FP32])
def test_autotuner_picks_grid_tile_and_waves(prec):
    s = systems.water_box(2000, 39.1, 5)
    print(f"case autotune prec={prec}", flush=True)
    ref = engine.MdSystem(s, precision=prec)
    _, ier = ref.runmd(20, update=10)
    assert ier == 0
    xr, vr = ref.get_state()
    ref.close()
    md = engine.MdSystem(s, precision=prec)
    md.set_option(native.JP_OPT_AUTOTUNE, 1)
    _, ier = md.runmd(20, update=10)
    assert ier == 0
    assert md.get_option(native.JP_OPT_GRID_SHIFT) in (-1, 0, 1)
    assert md.get_option(native.JP_OPT_LIST_BATCH) in (1, 2, 4)
    assert md.get_option(native.JP_OPT_LIST_WAVES) in (1, 2, 4, 8)
    assert md.get_option(native.JP_OPT_PAIR_WAVES) in (2, 4, 8)
    x, v = md.get_state()
    md.close()
    assert np.abs(x - xr).max() < (1e-9 if prec == engine.FP64 else 1e-5)
