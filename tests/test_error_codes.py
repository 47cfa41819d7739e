"""md_run's error returns on inputs that trigger them, from the oracle (CPU)
and through the C ABI (GPU), on the same coordinates:

  ier 28  celpar.f90:176-204: more atoms of one cluster class in a cell than
          ipmax (a crowded cell), or an atom outside the cell grid (a
          non-finite coordinate);
  ier 51  cutoff.f90:25-46: an atom with more partners in one list table
          than n15mx (a box at three times water's density);
  ier 1   shake.f90:286-319: SHAKE not converged within LIMSHK iterations
          (LIMSHK= 1, COVSHK= 1e-12);
  ier -2  shake.f90:266-284: a singular SHAKE system (a bond of zero length
          in the reference coordinates).
Every case is an ordinary return from a host-checked status word; no kernel
fails."""
import numpy as np
import pytest

from j_presto_amd import engine, systems
from oracle import Oracle


def dense_box():
    # 1,500 waters in a 24.8 A box: about 3100 partners per atom within
    # 13.5 A, against n15mx = 2172; about 90 atoms per md_run cell (ipmax 168)
    return systems.water_box(1500, 24.8, 41)


def crowded():
    # 60 waters squeezed into +-1 A about the centre of one md_run cell
    s = systems.water_box(500, 24.8, 7)
    x = s.x.copy()
    blk = x[:180] - x[:180].mean(0)
    x[:180] = np.array([15.775, 15.775, 15.775]) + blk / np.abs(blk).max()
    return s, x


def with_nan():
    s = systems.water_box(500, 24.8, 7)
    x = s.x.copy()
    x[5, 1] = np.nan
    return s, x


def shake_inputs():
    s = systems.water_box(500, 24.8, 7)
    rng = np.random.default_rng(1)
    x0 = s.x.copy()
    x1 = x0 + rng.normal(scale=0.05, size=x0.shape)
    xs = s.x.copy()
    xs[1] = xs[0]     # H1 on its O: the old bond vector is zero
    return s, x0, x1, xs


def test_oracle_error_codes():
    s = dense_box()
    assert Oracle(s).celprp(s.x) == 51
    s, x = crowded()
    assert Oracle(s).celprp(x) == 28
    s, x = with_nan()
    assert Oracle(s).celprp(x) == 28
    s, x0, x1, xs = shake_inputs()
    assert Oracle(s, shake_maxit=1, shake_tol=1e-12).shake(x1.copy(), x0) == 1
    assert Oracle(s).shake(x1.copy(), xs) == -2
    # the same inputs with a normal tolerance and geometry succeed
    assert Oracle(s).shake(x1.copy(), x0) == 0


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [engine.FP64, engine.FP32])
def test_error_codes_through_the_abi(prec):
    print(f"case error codes prec={prec}", flush=True)
    s = dense_box()
    md = engine.MdSystem(s, precision=prec)
    assert md.celprp() == 51
    md.close()
    s, x = crowded()
    md = engine.MdSystem(s, precision=prec)
    assert md.celprp(x) == 28
    md.close()
    s, x = with_nan()
    md = engine.MdSystem(s, precision=prec)
    assert md.celprp(x) == 28
    md.close()
    s, x0, x1, xs = shake_inputs()
    md = engine.MdSystem(s, precision=prec, shake_maxit=1, shake_tol=1e-12)
    assert md.shake(x1.copy(), x0) == 1
    md.close()
    md = engine.MdSystem(s, precision=prec)
    assert md.shake(x1.copy(), xs) == -2
    xg = x1.copy()
    assert md.shake(xg, x0) == 0
    md.close()
    # a normal system after all of it: no error left behind
    s = systems.water_box(500, 24.8, 7)
    md = engine.MdSystem(s, precision=prec)
    assert md.celprp() == 0
    _, ier = md.runmd(10, update=10)
    assert ier == 0
    md.close()
