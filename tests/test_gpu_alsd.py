"""Adaptive lambda square dynamics (md_run METHOD= GE under GEPSME= LMD,
verlethe.f90:235-290) on the GPU through the C ABI (jp_set_alsd): lambda
and its velocity live on the device, its force comes from the step's class
energies formed there, and the canonical kernels carry its kinetic term.

Bars: FP64 trajectories (x, v) within 1e-9 of md_run's after its restart
step, lambda and its velocity within 1e-11 (the group sums' order differs
from md_run's serial sum); graphs equal host launches bitwise; FP32/fixed
point within 1e-5 A and 1e-6 in lambda of the oracle."""
import numpy as np
import pytest

from conftest import alsd_params
from j_presto_amd import engine
from oracle import Oracle

pytestmark = pytest.mark.gpu


def run_gpu(s, f, tag, prec=engine.FP64, graphs=True, nmon=0):
    al, tempco = alsd_params(s, f, tag)
    md = engine.MdSystem(s, precision=prec)
    md.set_graphs(graphs)
    md.set_state(s.x, s.v)
    md.set_canonical(300.0, 3, 1e-3, tempco=tempco)
    md.set_alsd(al)
    md.set_lambda(s.lam, 0.0)
    assert md.celprp() == 0
    _, ier = md.restart_step()
    assert ier == 0
    assert md.get_lambda() == (s.lam, 0.0)   # the set-up step puts lambda back
    log, ier = md.runmd(int(f[f"alsd_{tag}_loop"]), update=10, nmon=nmon)
    assert ier == 0
    x, v = md.get_state()
    lam, lamv = md.get_lambda()
    md.close()
    return x, v, lam, lamv, log


@pytest.mark.parametrize("tag", ["a", "b", "c", "d"])
def test_alsd_vs_reference(golden, tag):
    """The four md_run ALSD runs on chain_lmd (windows; lower wall with
    LTCNTL= YES and TEMPCO= YES; upper wall without lambda's temperature
    control; TEMPCO= YES without it), FP64 with step graphs."""
    s, f = golden("chain_lmd")
    print(f"case alsd {tag}", flush=True)
    x, v, lam, lamv, _ = run_gpu(s, f, tag)
    print(f"max |dx| {np.abs(x - f[f'alsd_{tag}_x']).max():.2e} dlam {lam - float(f[f'alsd_{tag}_lam']):.2e}",
          flush=True)
    assert np.abs(x - f[f"alsd_{tag}_x"]).max() < 1e-9
    assert np.abs(v - f[f"alsd_{tag}_v"]).max() < 1e-9
    assert abs(lam - float(f[f"alsd_{tag}_lam"])) < 1e-11
    assert abs(lamv - float(f[f"alsd_{tag}_lamv"])) < 1e-11


def test_alsd_graphs_equal_eager_and_log(golden):
    """Step graphs replay the host-launched steps bit for bit (lambda's state
    included); with an energy record every step, lambda after each step
    matches md_run's NAMETO column to its printed digits and the record's
    E1, E2 those of the NAMETO file."""
    s, f = golden("chain_lmd")
    print("case alsd graphs/eager", flush=True)
    a = run_gpu(s, f, "a", graphs=True)
    b = run_gpu(s, f, "a", graphs=False, nmon=1)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert a[2] == b[2] and a[3] == b[3]
    log = b[4]
    ene = f["alsd_a_ene"]
    assert len(log) == len(ene)
    assert np.allclose(log["lambda"], ene[:, 0], rtol=1e-7, atol=0)
    e1 = log["ebond"][:, 0] + log["eangle"][:, 0] + log["etors"][:, 0] + log["eimp"][:, 0] + \
        log["evdw14"][:, 0] + log["eelec14"][:, 0] + log["evdw"][:, 0] + log["eelec"][:, 0]
    e2 = log["ebond"][:, 1] + log["eangle"][:, 1] + log["etors"][:, 1] + log["eimp"][:, 1] + \
        log["evdw14"][:, 1] + log["eelec14"][:, 1] + log["evdw"][:, 1] + log["eelec"][:, 1]
    assert np.allclose(e1, ene[:, 1], rtol=1e-6, atol=0)
    assert np.allclose(e2, ene[:, 2], rtol=1e-6, atol=0)


@pytest.mark.parametrize("tag", ["a", "b"])
def test_alsd_fp32_vs_oracle(golden, tag):
    """FP32 pair arithmetic with fixed-point sums: x within 1e-5 A and
    lambda within 1e-6 of the oracle's FP64 ALSD run."""
    s, f = golden("chain_lmd")
    print(f"case alsd fp32 {tag}", flush=True)
    al, tempco = alsd_params(s, f, tag)
    x, v, lam, lamv, _ = run_gpu(s, f, tag, prec=engine.FP32)
    o = Oracle(s)
    xo, vo = s.x.copy(), s.v.copy()
    ier, lam_o, lamv_o, _, _ = o.run_alsd(xo, vo, 0.0, int(f[f"alsd_{tag}_loop"]), al,
                                          grp=s.temperature_groups() if tempco else None)
    assert ier == 0
    assert np.abs(x - xo).max() < 1e-5
    assert abs(lam - lam_o) < 1e-6


def test_alsd_needs_lmd_and_canonical(golden):
    s, f = golden("chain_lmd")
    al, _ = alsd_params(s, f, "a")
    md = engine.MdSystem(s, precision=engine.FP64)
    with pytest.raises(Exception):
        md.set_alsd(al)            # no METHOD= CANO set-up
    md.set_canonical(300.0)
    md.set_alsd(al)
    md.set_alsd(None)
    md.close()
    w, _ = golden("water_1500")
    md = engine.MdSystem(w, precision=engine.FP64)
    md.set_canonical(300.0)
    with pytest.raises(Exception):
        md.set_alsd(al)            # one energy class
    md.close()
