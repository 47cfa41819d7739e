"""The oracle's cell multipole method (md_run CAL15M= CMM, BOUNDA= NO) against
the compiled reference md_run (fixtures droplet_cmm, chain_droplet_cmm,
chain_droplet_lmd made by tests/golden/make_golden.py --cmm): the
FDEBUG= MAKE force dump bit for bit and MICR trajectories within 1e-9 A,
where the far field is recomputed every CMMUPD= 5 steps and reused in
between, and the tables and multipoles every UPDATE= 10 steps.  Beside
them two checks that fix the method by mathematics on an ion lattice: the
direct Coulomb sum when every charge sits on its cell's centre, and an
error of third order in the charges' offsets otherwise."""
import numpy as np
import pytest

from oracle import Oracle

CMM_NAMES = ["droplet_cmm", "chain_droplet_cmm", "chain_droplet_lmd"]


def cmm_forces(s, itwin=True):
    o = Oracle(s)
    assert o.cmm_setup(s.x) == 0
    assert o.cmm_celprp(s.x) == 0
    g, et = o.bonded(s.x)
    g, ee, ev = o.cmm_nonbond(s.x, g, itwin=itwin)
    return o, o.combine(g), et, ee, ev


@pytest.mark.parametrize("name", CMM_NAMES)
def test_cmm_forces_bit_exact(golden, name):
    s, f = golden(name)
    _, gt, et, ee, ev = cmm_forces(s)
    print(name, s.n, "max |dg|", np.abs(gt - f["force_grad"]).max())
    assert np.array_equal(gt, f["force_grad"])
    # dnergy slots 8 VDW15 and 9 ELE15 hold the near and far sums of every class
    assert ee.sum() == f["force_epot"][8]
    assert ev.sum() == f["force_epot"][7]
    for k in range(6):
        assert et[:, k].sum() == f["force_epot"][1 + k], k


@pytest.mark.parametrize("name", CMM_NAMES)
def test_cmm_micr_trajectory(golden, name):
    """md_run's MICR loop with SHAKE, UPDATE= 10 and CMMUPD= 5 (the far field
    reused on 4 of 5 steps)."""
    s, f = golden(name)
    o = Oracle(s)
    assert o.cmm_setup(s.x) == 0
    x, v = s.x.copy(), s.v.copy()
    n = int(f["traj_nsteps"])
    assert o.cmm_run_nve(x, v, n, 10, 5) == 0
    dx, dv = np.abs(x - f["traj_x"]).max(), np.abs(v - f["traj_v"]).max()
    print(name, n, "steps: max |dx|", dx, "max |dv|", dv)
    assert dx < 1e-9 and dv < 1e-9


def test_cmm_far_field_reuse_differs_from_fresh(golden):
    """The far-field reuse is real: with itwin off the stored far field of
    the set-up call is added, so moving the atoms and reusing it gives other
    forces than a fresh far field, and at the stored positions the same."""
    s, _ = golden("droplet_cmm")
    o = Oracle(s)
    assert o.cmm_setup(s.x) == 0 and o.cmm_celprp(s.x) == 0
    g1, e1, _ = o.cmm_nonbond(s.x, itwin=True)
    g2, e2, _ = o.cmm_nonbond(s.x, itwin=False)
    assert np.array_equal(g1, g2) and np.array_equal(e1, e2)
    x2 = s.x + np.random.default_rng(3).normal(scale=0.05, size=s.x.shape)
    g3, _, _ = o.cmm_nonbond(x2, itwin=False)
    g4, _, _ = o.cmm_nonbond(x2, itwin=True)
    assert not np.array_equal(g3, g4)
    assert np.abs(g3 - g4).max() < 0.2 * np.abs(g4).max()


def ion_lattice(delta: float, seed: int = 5):
    """One +1 or -1 ion (no LJ) at every centre of a 4x4x4 grid of 6 A cells
    (+-3, +-9 A), each moved by delta along a fixed random direction: one
    CMM level, whose far field is every cell beyond the 27 nearest."""
    from j_presto_amd import systems
    rng = np.random.default_rng(seed)
    g = np.array([-9.0, -3.0, 3.0, 9.0])
    x = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    u = rng.normal(size=x.shape)
    u /= np.linalg.norm(u, axis=1)[:, None]
    q = np.where(rng.random(len(x)) < 0.5, 1.0, -1.0)
    pos, neg = np.where(q > 0)[0], np.where(q < 0)[0]
    order = np.concatenate([pos, neg])
    mols = [systems.ion_template("P", len(pos), 1, 23.0, 1.0), systems.ion_template("N", len(neg), 1, 35.0, -1.0)]
    s = systems.System(name="ions", mols=mols, types=[systems.AtomType("ION", 1.5, 0.0)],
                       box=np.array([40.0, 40.0, 40.0]), x=(x + delta * u)[order].copy(),
                       v=np.zeros((len(x), 3)))
    return s


def direct_coulomb(s, ccoef):
    d = s.x[:, None, :] - s.x[None, :, :]
    r2 = (d * d).sum(-1)
    np.fill_diagonal(r2, np.inf)
    q = s.charge
    e = ccoef * q[:, None] * q[None, :] / np.sqrt(r2)
    return (-(e / r2)[:, :, None] * d).sum(1), 0.5 * e.sum()


def cmm_error(delta):
    s = ion_lattice(delta)
    o = Oracle(s)
    assert o.cmm_setup(s.x, 6.0, 1) == 0 and o.cmm_celprp(s.x) == 0
    g, ee, ev = o.cmm_nonbond(s.x)
    gd, ed = direct_coulomb(s, o.constants()["ccoef"])
    assert ev.sum() == 0.0
    return np.abs(g.sum(0) - gd).max(), abs(ee.sum() - ed), np.abs(gd).max()


def test_cmm_exact_for_charges_at_cell_centres():
    """Charges sitting on their cells' centres have no dipole or quadrupole:
    near table plus far field is then the direct Coulomb sum."""
    eg, ee, scale = cmm_error(0.0)
    print("at centres: max |g - direct|", eg, "of", scale, "|dE|", ee)
    assert eg < 1e-12 * scale and ee < 1e-10


def test_cmm_far_field_third_order():
    """Moved off the centres by delta, the multipole expansion to the
    quadrupole leaves an error of order delta^3: halving delta divides the
    force error by about 8 (a wrong dipole would leave order delta, a wrong
    quadrupole order delta^2, i.e. ratios of 2 and 4)."""
    e1 = cmm_error(0.4)[0]
    e2 = cmm_error(0.2)[0]
    e3 = cmm_error(0.1)[0]
    print("errors", e1, e2, e3, "ratios", e1 / e2, e2 / e3)
    assert e1 / e2 > 6.5 and e2 / e3 > 6.5
