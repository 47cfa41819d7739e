import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests", "golden")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")
GOLDEN_NAMES = ["water_1500", "water_6k", "chain_3k", "chain_lmd", "ions_3k", "water_54", "chain_bonded",
                "chain_lmd_bonded", "chain_tors", "chain_lmd_tors", "water_42", "methyl_water"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")


def load_golden(name):
    """(System rebuilt from the fixture's recipe, fixture arrays).  The
    generator must reproduce the exact inputs md_run saw (fingerprint)."""
    from make_golden import build_system, fingerprint
    f = dict(np.load(os.path.join(GOLDEN, f"{name}.npz")))
    rec = json.loads(str(f["recipe"]))
    s = build_system(rec["kind"], rec["args"])
    assert fingerprint(s) == str(f["fingerprint"]), f"generator no longer reproduces {name}; regenerate fixtures"
    return s, f


@pytest.fixture(scope="session")
def golden():
    cache = {}

    def get(name):
        if name not in cache:
            cache[name] = load_golden(name)
        return cache[name]
    return get


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def oracle_monitor(o, s, x, v, nsteps, every=5, cano=None, scaled=False):
    """Runs the oracle's md_run loop (MICR, or CANO with cano=(temps, grp))
    and returns rows (step, E_tot, E_pot, K) at every `every`-th step:
    K = caltmp of (v(t-dt/2)+v(t+dt/2))/2 (md.f90:618-628), E_pot md_run's
    fdepot(1), the unscaled sum of the class energies (dnergy.f90:78-79;
    with scaled=True the lambda-scaled sum the LMD dynamics conserve,
    verlet.f90:67-91), E_tot = K + E_pot (md.f90:805, monmd.f90:123)."""
    lam = s.lam
    fac = [lam * lam, lam, 1.0] if (o.nfrg == 3 and scaled) else [1.0] * o.nfrg
    rows = []
    assert o.celprp(x) == 0
    for idloop in range(1, nsteps + 1):
        g, et = o.bonded(x)
        g, ee, ev = o.nonbond(x, g)
        ep = sum(f * (ee[k] + ev[k] + et[k].sum()) for k, f in enumerate(fac))
        if cano is None:
            vb = v.copy()
            assert o.verlet_step(x, v, o.combine(g)) == 0
        else:
            ier, vb = o.cano_step(x, v, o.combine(g), cano[0], 3, 1e-3, cano[1])
            assert ier == 0
        if idloop % every == 0:
            ek = o.ekin(0.5 * (vb + v))
            rows.append((idloop, ek + ep, ep, ek))
        if idloop % 10 == 0:
            assert o.celprp(x) == 0
    return np.array(rows)


def ge_params(f, tag):
    """The multicanonical parameters of a fixture's METHOD= GE run `tag`."""
    celw, ceup = f[f"ge_{tag}_range"]
    return dict(c=f["ge_poly"], low=f["ge_low"], high=f["ge_high"], alphalw=float(f["ge_alpha"][0]),
                alphaup=float(f["ge_alpha"][1]), celw=float(celw), ceup=float(ceup), tempce=float(f["ge_tempce"]))


def md_cell_windows(L, nv, rc=12.0):
    """md_run's periodic neighbour-cell table per axis (incel.f90:334-358,
    456-508), restated from the reference for the tests: {cell: set of cells
    in its window} for each edge in L; None where md_run writes out of bounds."""
    import math
    rlim = rc + 1.5
    siz = 0.5 * rlim
    m = math.ceil(rlim / siz)
    ic = int(math.floor(nv * 0.5 + 0.5))
    out = []
    for edge in L:
        h = 0.5 * edge / siz
        m1, m2 = math.ceil(ic - h), math.ceil(ic + h)
        mdel = 1 if math.ceil(h) - h > 0.5 else 2
        if m1 < 1 or m2 > nv:
            return None
        ii = m2 - m1 + 1
        ri = np.float32(1.0) / np.float32(ii)
        win = {}
        for c in range(m1, m2 + 1):
            lo, hi = c - m, c + m
            if lo <= m1:
                lo -= mdel
            if hi >= m2:
                hi += mdel
            w = set()
            for k in range(lo, hi + 1):
                kk = k - ii * int(math.floor(np.float32(k - m1) * ri))
                if kk < m1 or kk > m2:
                    return None
                w.add(kk)
            win[c] = w
        out.append(win)
    return out


def admissible(windows, ci, cj):
    """Pairs (i, j) whose md_run cells are in md_run's table: cj's index in
    ci's window on every axis (ci, cj: (n, 3) 1-based cells)."""
    ok = np.ones(len(ci), bool)
    for a in range(3):
        w = windows[a]
        ok &= np.array([int(b) in w[int(c)] for c, b in zip(ci[:, a], cj[:, a])], bool)
    return ok


def alsd_params(s, f, tag):
    """The ALSD parameters of a fixture's run `tag` (make_golden.ALSD_*):
    lambda's mass is LWEIGH, or md_run's default, the class-1 atoms' masses
    added in atom order (inpcls.f90:348-357)."""
    celw, ceup = f[f"alsd_{tag}_range"]
    tempco, tctl, lw = f[f"alsd_{tag}_opt"]
    lm = float(lw)
    if lm == 0.0:
        for m, c in zip(s.mass, s.cluster):
            if c == 1:
                lm += float(m)
    return dict(c=f["alsd_poly"], low=f["alsd_low"], high=f["alsd_high"], lower=float(f["alsd_walls"][0]),
                upper=float(f["alsd_walls"][1]), celw=float(celw), ceup=float(ceup),
                tempce=float(f["alsd_tempce"]), lmass=lm, tctl=bool(tctl)), bool(tempco)
