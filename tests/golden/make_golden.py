"""Writes the golden fixtures under tests/golden/ by running the compiled
reference md_run (oracle/_ref/md_run, built from /root/reference by
oracle/build_ref.sh) on synthetic systems.

For every system two md_run jobs run on one OpenMP thread with PARALV= DOUB
and N_SIMD= 1 (a fixed summation order):
  force job  RESTAR= YES, METHOD= CANO, FDEBUG= MAKE: md_run evaluates dnergy
             at the restart coordinates and dumps epot(15) and grad(3,N)
             (verlethe.f90:378-384) -- the FP64 gradient and energies at
             exactly the stored coordinates;
  traj job   RESTAR= YES, METHOD= MICR, LOOPLI= K: one restart step plus K
             leap-frog steps with SHAKE; the output restart holds x and v in
             FP64 and NAMETO the potential energy of every step.
  cano job   (some systems) RESTAR= YES, METHOD= CANO, SETTEM= 300,
             TEMPCO= NO or YES, LOOPLI= K: K velocity-scaled steps
             (verlethe.f90); the output restart holds x and v.
  mon jobs   (some systems) the traj and cano jobs again with OUTLOG= 5:
             md_run's monitoring blocks (monmd.f90 outtlg) print, every 5th
             step, the loop number, E_tot, E_pot, T and the kinetic energy
             caltmp computes from (v(t-dt/2)+v(t+dt/2))/2 (md.f90:618-628),
             each to 8 significant digits (T: 7 decimals).
  ge jobs    (water_1500, --ge-only) RESTAR= YES, METHOD= GE with a NAMEGE=
             parameter file (GE_POLY etc. below), SETTEM= 300, TEMPCO= NO:
             multicanonical steps whose forces carry the weight of the
             class-1 energy (verlethe.f90:196-233).
A fixture stores the generator recipe, the input x and v, and the outputs.
The reference itself never leaves this container.

  heat jobs  (--heat-only: HEAT_RUNS) METHOD= CANO through md_run's heating loop
             (HEATLO= 10, STARTT= 100), from restarts at loop 0 and 4.
  stop jobs  (--stop-only name ...: STOP_RUNS) STOPCE= TRAN, ROTA and BOTH under
             METHOD= MICR and CANO, and runs without a restart (RESTAR= NO,
             INITIA= ZERO: md_run's initial half step, staver).
  cmm jobs   (--cmm: CMM_RECIPES) CAL15M= CMM, BOUNDA= NO on droplets: the
             FDEBUG= MAKE dump and a MICR run with NAMETO.
  alsd jobs  (chain_lmd, --alsd-only) RESTAR= YES, METHOD= GE with GEPSME= LMD
             (adaptive lambda square dynamics, ALSD_* below): lambda moves
             under -2 E1 lambda - E2 - dew(lambda) beside the atoms.

Usage: python tests/golden/make_golden.py [--cano-only | --mon-only | --host-only | --ge-only | --alsd-only |
       --heat-only | --stop-only | --cmm]
       [name ...]
"""
from __future__ import annotations

import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from j_presto_amd import refio, systems  # noqa: E402

MD_RUN = os.path.join(ROOT, "oracle", "_ref", "md_run")
HERE = os.path.dirname(os.path.abspath(__file__))

RECIPES = {
    # small box just above 2*cutoff (box < 2*rlim: periodic images inside the list radius)
    "water_1500": ("water_box", dict(nwat=500, L=24.8, seed=7),
                   dict(traj=20, cano=("NO", 20), mon=True, host=True)),
    # the 6k-atom water workload itself
    "water_6k": ("workload", dict(i=0), dict(traj=20)),
    # chain with 16 LJ types, 2-atom SHAKE groups, chain exclusions
    "chain_3k": ("chain_in_water", dict(n_chain=200, L=30.0, seed=11), dict(traj=20, cano=("YES", 20))),
    # LMD: three energy classes, lambda-scaled gradient
    "chain_lmd": ("chain_in_water", dict(n_chain=160, L=28.0, seed=12, lmd=0.7),
                  dict(traj=10, cano=("YES", 10), mon=True)),
    # chain with harmonic bonds and angles (dnebnd_PB / dneang_PB), strained off their minima
    "chain_bonded": ("chain_in_water", dict(n_chain=200, L=30.0, seed=15, bond_fc=300.0, angle_fc=50.0,
                                            bond_r0=1.53, angle_deg=112.0), dict(traj=20)),
    # LMD with bonded terms: bonds and angles in the unscaled class (SCALED_TERM default)
    "chain_lmd_bonded": ("chain_in_water", dict(n_chain=120, L=26.0, seed=16, lmd=0.7, bond_fc=300.0,
                                                angle_fc=50.0, bond_r0=1.53, angle_deg=112.0), dict(traj=10)),
    # chain with torsions and 1-4 terms on top of bonds and angles (dnei14_PB)
    "chain_tors": ("chain_in_water", dict(n_chain=200, L=30.0, seed=17, bond_fc=300.0, angle_fc=50.0,
                                          bond_r0=1.53, angle_deg=112.0, torsion_fc=1.0), dict(traj=20)),
    # LMD with torsions: torsion classes from the middle atoms, 1-4 classes from the pair
    "chain_lmd_tors": ("chain_in_water", dict(n_chain=120, L=26.0, seed=18, lmd=0.7, bond_fc=300.0,
                                              angle_fc=50.0, bond_r0=1.53, angle_deg=112.0, torsion_fc=1.0),
                       dict(traj=10)),
    # ions with their own LJ types
    "ions_3k": ("water_box", dict(nwat=900, L=30.0, seed=13, n_ion_pairs=10), dict(traj=0)),
    # box edge a multiple of rlim = 13.5 A (md_run's phantom-cell case)
    "water_54": ("water_box", dict(nwat=5200, L=54.0, seed=14), dict(traj=0)),
    # rigid CH3 groups in water: 4-atom SHAKE groups (shake4_PB)
    "methyl_water": ("methyl_water", dict(nmet=40, nwat=460, L=30.0, seed=31), dict(traj=20, cano=("NO", 12))),
    # 42 A edge: md_run's periodic cell table (8 cells of 6.75 A for 42 A) never tests
    # pairs between some cells facing each other across the boundary, inside the cutoff
    "water_42": ("water_box", dict(nwat=2469, L=42.0, seed=19), dict(traj=10)),
}


# Non-periodic systems under CAL15M= CMM (BOUNDA= NO): cmm job = force dump
# plus a MICR run of `traj` steps (UPDATE= 10, CMMUPD= 5 by default: the far
# field is recomputed every 5th step and reused in between, the multipoles and
# tables every 10th).
CMM_RECIPES = {
    # 481-water droplet, 15 A radius: 16^3 level-1 cells of 6 A, three levels
    "droplet_cmm": ("water_droplet", dict(R=15.0, seed=21), dict(traj=20, cano=12, mon=200)),
    # compact 120-atom chain with bonds, angles, torsions and 1-4 terms in a droplet
    "chain_droplet_cmm": ("chain_droplet", dict(n_chain=120, R=14.0, seed=41, bond_fc=300.0, angle_fc=50.0,
                                                torsion_fc=1.0), dict(traj=20)),
    # the same under LMD (two fragments, three energy classes)
    "chain_droplet_lmd": ("chain_droplet", dict(n_chain=120, R=14.0, seed=41, lmd=0.7, bond_fc=300.0,
                                                angle_fc=50.0, torsion_fc=1.0), dict(traj=10, cano=12, alsd="a")),
}


def build_system(kind: str, args: dict) -> systems.System:
    if kind == "workload":
        return systems.workload(args["i"])
    return getattr(systems, kind)(**args)


def fingerprint(s: systems.System) -> str:
    h = hashlib.sha256()
    for a in (s.x, s.v, s.charge, s.mass, s.atype.astype(np.int64)):
        h.update(np.ascontiguousarray(a).tobytes())
    h.update(json.dumps([(t.rmin2, t.eps) for t in s.types]).encode())
    return h.hexdigest()


def run_md(workdir: str, s: systems.System, md: dict, shake: bool = True, inp: dict = None) -> str:
    files = refio.write_inputs(workdir, s, shake=shake)
    refio.write_md_inp(os.path.join(workdir, "md.inp"), files, s, md, inp=inp)
    env = dict(os.environ, OMP_NUM_THREADS="1", OMP_STACKSIZE="512M")
    # md_run keeps per-atom automatic arrays on the stack: lift the limit for large systems
    with open(os.path.join(workdir, "md.inp")) as fin, open(os.path.join(workdir, "md.log"), "w") as fout:
        subprocess.run(["bash", "-c", f"ulimit -s unlimited; exec {MD_RUN}"], stdin=fin, stdout=fout,
                       stderr=subprocess.STDOUT, cwd=workdir, env=env, check=True, timeout=3600)
    return open(os.path.join(workdir, "md.log")).read()


def make(name: str) -> str:
    kind, args, opt = RECIPES[name]
    s = build_system(kind, args)
    out = dict(recipe=json.dumps(dict(kind=kind, args=args)), fingerprint=fingerprint(s),
               x=s.x, v=s.v, box=s.box)
    tmp = tempfile.mkdtemp(prefix="jpgold_")
    try:
        refio.write_restart(os.path.join(tmp, "in.rst"), s.x, s.v, lam=s.lam)
        log = run_md(tmp, s, {"METHOD": "CANO", "RESTAR": "YES", "NAMERI": "in.rst", "NAMERO": "out.rst",
                              "LOOPLI": "1", "FDEBUG": "MAKE", "SETTEM": "300.0"})
        dump = refio.read_force_dump(os.path.join(tmp, "1stp_force.bin"), s.n)
        out["force_grad"] = dump["grad"]
        out["force_epot"] = dump["epot"]
        ntraj = opt.get("traj", 0)
        if ntraj:
            t2 = os.path.join(tmp, "traj")
            os.makedirs(t2)
            refio.write_restart(os.path.join(t2, "in.rst"), s.x, s.v, lam=s.lam)
            log = run_md(t2, s, {"METHOD": "MICR", "RESTAR": "YES", "NAMERI": "in.rst", "NAMERO": "out.rst",
                                 "LOOPLI": str(ntraj), "UPDATE": "10", "TIMEST": "2.0",
                                 "NAMETO": "ene.dat", "OUTLOG": "100000",
                                 # flang evaluates both operands of `idmoce.gt.0 .and. mod(idloop,idmoce)`
                                 # (md.f90:604): a positive OUTCLE keeps mod() away from zero
                                 "OUTCLE": "100000"})
            r = refio.read_restart(os.path.join(t2, "out.rst"))
            out["traj_nsteps"] = np.int64(ntraj)
            out["traj_x"] = r["x"]
            out["traj_v"] = r["v"]
            out["traj_loop"] = np.int64(r["loop"])
            ene = np.loadtxt(os.path.join(t2, "ene.dat"), ndmin=2)
            out["traj_epot"] = ene
        out.update(cano_job(s, opt, tmp))
        if opt.get("mon"):
            out.update(mon_jobs(s, opt, tmp))
        if opt.get("host"):
            out.update(host_jobs(s, tmp))
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **out)
        return path
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def cano_job(s: systems.System, opt: dict, tmp: str) -> dict:
    if "cano" not in opt:
        return {}
    tempco, ncano = opt["cano"]
    t3 = os.path.join(tmp, "cano")
    os.makedirs(t3)
    refio.write_restart(os.path.join(t3, "in.rst"), s.x, s.v, lam=s.lam)
    run_md(t3, s, {"METHOD": "CANO", "RESTAR": "YES", "NAMERI": "in.rst", "NAMERO": "out.rst",
                   "LOOPLI": str(ncano), "UPDATE": "10", "TIMEST": "2.0", "SETTEM": "300.0",
                   "TEMPCO": tempco, "OUTLOG": "100000", "OUTCLE": "100000"})
    r = refio.read_restart(os.path.join(t3, "out.rst"))
    return dict(cano_tempco=np.array(tempco), cano_nsteps=np.int64(ncano), cano_x=r["x"], cano_v=r["v"])


MON_KEYS = ("MD LOOP NUMBER", "TOTAL      (KCAL/MOL)", "POTENTIAL  (KCAL/MOL)", "TEMPERATURE (K)",
            "KINETIC    (KCAL/MOL)")


def parse_monitor(log: str) -> np.ndarray:
    """(step, E_tot, E_pot, T, K) of every monitoring block md_run printed
    after its loop started (outtlg's format 810: 'NAME : value' fields, two
    per line); the restart's check block before the loop is skipped."""
    body = log[log.index("MOLECULAR DYNAMICS LOOP START"):]
    rows, cur = [], {}
    for line in body.splitlines():
        for key in MON_KEYS:
            k = line.find(key)
            if k < 0:
                continue
            rest = line[k + len(key):].lstrip()
            if not rest.startswith(":"):
                continue
            cur[key] = float(rest[1:].split()[0])
        if all(k in cur for k in MON_KEYS):
            rows.append([cur[k] for k in MON_KEYS])
            cur = {}
    return np.array(rows, np.float64)


def mon_jobs(s: systems.System, opt: dict, tmp: str) -> dict:
    """The traj job (and the cano job) again with OUTLOG= 5; returns the
    parsed monitoring rows."""
    out = {}
    common = {"RESTAR": "YES", "NAMERI": "in.rst", "NAMERO": "out.rst", "UPDATE": "10", "TIMEST": "2.0",
              "OUTLOG": "5", "OUTCLE": "100000"}
    d = os.path.join(tmp, "mon")
    os.makedirs(d)
    refio.write_restart(os.path.join(d, "in.rst"), s.x, s.v, lam=s.lam)
    log = run_md(d, s, dict(common, METHOD="MICR", LOOPLI=str(opt["traj"])))
    out["mon_micr"] = parse_monitor(log)
    if "cano" in opt:
        tempco, ncano = opt["cano"]
        d = os.path.join(tmp, "cmon")
        os.makedirs(d)
        refio.write_restart(os.path.join(d, "in.rst"), s.x, s.v, lam=s.lam)
        log = run_md(d, s, dict(common, METHOD="CANO", LOOPLI=str(ncano), SETTEM="300.0", TEMPCO=tempco))
        out["mon_cano"] = parse_monitor(log)
    return out


def host_jobs(s: systems.System, tmp: str) -> dict:
    """Two md_run runs for the md.inp host (jp_host --mdinp):
      rst25   a restart whose loop number is 25, LOOPLI= 20, UPDATE= 10,
              OUTLOG= 5: md_run's idloop runs 26..45, so the list is rebuilt
              after steps 30 and 40 and energies print at 30, 35, 40, 45;
      cputim  LOOPLI= 30 with CPUTIM= 0.0001 s: md_run stops at its first
              list update (loop 10) and writes its restart (md.f90:655-676);
      retry   METHOD= CANO with LIMSHK= 3: SHAKE does not converge within 3
              iterations at dt = 2 fs, so every step takes verlethe's retry
              (verlethe.f90:552-573: undo, halve dt, redo the forces)."""
    out = {}
    common = {"METHOD": "MICR", "RESTAR": "YES", "NAMERI": "in.rst", "NAMERO": "out.rst", "UPDATE": "10",
              "TIMEST": "2.0", "OUTCLE": "100000"}
    d = os.path.join(tmp, "rst25")
    os.makedirs(d)
    refio.write_restart(os.path.join(d, "in.rst"), s.x, s.v, loop=25, lam=s.lam)
    log = run_md(d, s, dict(common, LOOPLI="20", OUTLOG="5"))
    r = refio.read_restart(os.path.join(d, "out.rst"))
    out.update(rst25_x=r["x"], rst25_v=r["v"], rst25_loop=np.int64(r["loop"]), rst25_mon=parse_monitor(log))
    d = os.path.join(tmp, "retry")
    os.makedirs(d)
    refio.write_restart(os.path.join(d, "in.rst"), s.x, s.v, lam=s.lam)
    log = run_md(d, s, dict(common, METHOD="CANO", SETTEM="300.0", TEMPCO="NO", LOOPLI="20", OUTLOG="100000",
                            LIMSHK="3"))
    r = refio.read_restart(os.path.join(d, "out.rst"))
    out.update(retry_x=r["x"], retry_v=r["v"], retry_loop=np.int64(r["loop"]),
               retry_provisions=np.int64(log.count("SHAKE provision")))
    d = os.path.join(tmp, "cputim")
    os.makedirs(d)
    refio.write_restart(os.path.join(d, "in.rst"), s.x, s.v, lam=s.lam)
    run_md(d, s, dict(common, LOOPLI="30", OUTLOG="100000", CPUTIM="0.0001"))
    r = refio.read_restart(os.path.join(d, "out.rst"))
    out.update(cputim_x=r["x"], cputim_v=r["v"], cputim_loop=np.int64(r["loop"]))
    return out


# Multicanonical (METHOD= GE) parameter sets for water_1500: three windows of
# cubic polynomials in the reduced energy, weights near 1 at 300 K; the
# energy range [celw, ceup] puts the run in the windows ("win"), below celw
# ("lo": alphalw) or above ceup ("hi": alphaup) for its whole length.
GE_POLY = np.array([[1.72, -0.20, 0.15, -0.05], [1.68, -0.12, 0.10, 0.03], [1.64, 0.05, -0.08, 0.02]])
GE_LOW, GE_HIGH = np.array([0.0, 0.3, 0.6]), np.array([0.3, 0.6, 1.0])
GE_RUNS = {"win": (-4500.0, 0.0, 20), "lo": (1000.0, 2000.0, 10), "hi": (-9000.0, -8000.0, 10)}
GE_ALPHA = (1.70, 1.62)      # alphalw, alphaup
GE_TEMPCE = 300.0


def write_ge_params(path: str, celw: float, ceup: float) -> None:
    """inmcmd.f90:20-31 layout: nwindow ndeg; per window one coefficient per
    record, then low_v high_v; alphalw alphaup; celw ceup; tempce."""
    nw, nd1 = GE_POLY.shape
    L = [f"{nw} {nd1 - 1}"]
    for w in range(nw):
        L += [repr(float(c)) for c in GE_POLY[w]]
        L.append(f"{float(GE_LOW[w])!r} {float(GE_HIGH[w])!r}")
    L += [f"{GE_ALPHA[0]!r} {GE_ALPHA[1]!r}", f"{celw!r} {ceup!r}", repr(GE_TEMPCE)]
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def ge_jobs(s: systems.System, tmp: str) -> dict:
    """METHOD= GE (idfvsl = 4, verlethe.f90:196-233) runs from the fixture's
    restart with TEMPCO= NO at 300 K, one per GE_RUNS entry: final x, v."""
    out = dict(ge_poly=GE_POLY, ge_low=GE_LOW, ge_high=GE_HIGH, ge_alpha=np.array(GE_ALPHA),
               ge_tempce=np.array(GE_TEMPCE))
    for tag, (celw, ceup, nsteps) in GE_RUNS.items():
        d = os.path.join(tmp, "ge_" + tag)
        os.makedirs(d)
        refio.write_restart(os.path.join(d, "in.rst"), s.x, s.v, lam=s.lam)
        write_ge_params(os.path.join(d, "ge.par"), celw, ceup)
        run_md(d, s, {"METHOD": "GE", "RESTAR": "YES", "NAMERI": "in.rst", "NAMERO": "out.rst", "UPDATE": "10",
                      "TIMEST": "2.0", "SETTEM": "300.0", "TEMPCO": "NO", "NAMEGE": "ge.par",
                      "LOOPLI": str(nsteps), "OUTLOG": "100000", "OUTCLE": "100000"})
        r = refio.read_restart(os.path.join(d, "out.rst"))
        out.update({f"ge_{tag}_x": r["x"], f"ge_{tag}_v": r["v"], f"ge_{tag}_loop": np.int64(r["loop"]),
                    f"ge_{tag}_range": np.array([celw, ceup])})
    return out


# Adaptive lambda square dynamics (METHOD= GE under GEPSME= LMD) for
# chain_lmd: two windows of cubic polynomials in lambda itself (bias dew of
# several hundred kcal/mol, of the order of -2 E1 lambda - E2 at lambda =
# 0.7), walls lower/upper outside [celw, ceup].  Runs (tag: celw, ceup,
# TEMPCO, LTCNTL, LWEIGH or None for md_run's default, the class-1 mass):
# inside the windows, below celw, above ceup, with and without lambda's
# temperature control, one and two temperature groups.
ALSD_POLY = np.array([[1400.0, 300.0, -200.0, 100.0], [1350.0, 400.0, -150.0, 80.0]])
ALSD_LOW, ALSD_HIGH = np.array([0.0, 0.65]), np.array([0.65, 1.5])
ALSD_WALLS = (2000.0, 3000.0)   # lower, upper
ALSD_TEMPCE = 300.0
ALSD_RUNS = {"a": (0.5, 1.0, "NO", "YES", None), "b": (0.8, 0.95, "YES", "YES", 400.0),
             "c": (0.4, 0.6, "NO", "NO", None), "d": (0.5, 1.0, "YES", "NO", 250.0)}


def write_alsd_params(path: str, celw: float, ceup: float) -> None:
    """inmcmd.f90:20-41 layout under LMD: the GE records, then lower upper."""
    nw, nd1 = ALSD_POLY.shape
    L = [f"{nw} {nd1 - 1}"]
    for w in range(nw):
        L += [repr(float(c)) for c in ALSD_POLY[w]]
        L.append(f"{float(ALSD_LOW[w])!r} {float(ALSD_HIGH[w])!r}")
    L += ["1.0 1.0", f"{celw!r} {ceup!r}", repr(ALSD_TEMPCE), f"{ALSD_WALLS[0]!r} {ALSD_WALLS[1]!r}"]
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def alsd_jobs(s: systems.System, tmp: str) -> dict:
    """ALSD runs (verlethe.f90 idfvsl = 4 under LMD) from the fixture's
    restart at 300 K, 10 steps each: final x, v, lambda, lambda_v and the
    NAMETO series (lambda after the step, E1, E2, their sum with the
    unscaled class 3, md.f90:578-583); run "a" also prints its monitoring
    blocks every 5 steps (kinetic energy with lambda's term)."""
    out = dict(alsd_poly=ALSD_POLY, alsd_low=ALSD_LOW, alsd_high=ALSD_HIGH, alsd_walls=np.array(ALSD_WALLS),
               alsd_tempce=np.array(ALSD_TEMPCE))
    for tag, (celw, ceup, tempco, ltc, lw) in ALSD_RUNS.items():
        d = os.path.join(tmp, "alsd_" + tag)
        os.makedirs(d)
        refio.write_restart(os.path.join(d, "in.rst"), s.x, s.v, lam=s.lam)
        write_alsd_params(os.path.join(d, "ge.par"), celw, ceup)
        inp = {"LTCNTL": ltc}
        if lw is not None:
            inp["LWEIGH"] = repr(float(lw))
        log = run_md(d, s, {"METHOD": "GE", "RESTAR": "YES", "NAMERI": "in.rst", "NAMERO": "out.rst",
                            "UPDATE": "10", "TIMEST": "2.0", "SETTEM": "300.0", "TEMPCO": tempco,
                            "NAMEGE": "ge.par", "NAMETO": "ene.dat", "LOOPLI": "10",
                            "OUTLOG": "5" if tag == "a" else "100000", "OUTCLE": "100000"}, inp=inp)
        r = refio.read_restart(os.path.join(d, "out.rst"))
        ene = np.loadtxt(os.path.join(d, "ene.dat"), ndmin=2)
        out.update({f"alsd_{tag}_x": r["x"], f"alsd_{tag}_v": r["v"], f"alsd_{tag}_lam": np.float64(r["lam"]),
                    f"alsd_{tag}_lamv": np.float64(r["lam_v"]), f"alsd_{tag}_loop": np.int64(r["loop"]),
                    f"alsd_{tag}_range": np.array([celw, ceup]), f"alsd_{tag}_ene": ene,
                    f"alsd_{tag}_opt": np.array([1 if tempco == "YES" else 0, 1 if ltc == "YES" else 0,
                                                 0.0 if lw is None else lw])})
        if tag == "a":
            out["alsd_a_mon"] = parse_monitor(log)
    return out


def make_cmm(name: str) -> str:
    """CAL15M= CMM, BOUNDA= NO fixture: md_run's FDEBUG= MAKE dump at the
    stored coordinates and a MICR run with its NAMETO series; with cano a
    METHOD= CANO run at 300 K (TEMPCO= NO), with alsd an ALSD run with the
    parameters of ALSD_RUNS[alsd], and with mon a MICR run of mon steps whose
    monitoring blocks (every 20 steps) are kept, each from md_run's restart
    set-up."""
    kind, args, opt = CMM_RECIPES[name]
    s = build_system(kind, args)
    out = dict(recipe=json.dumps(dict(kind=kind, args=args)), fingerprint=fingerprint(s), x=s.x, v=s.v)
    cmm = {"CAL15M": "CMM", "BOUNDA": "NO"}
    tmp = tempfile.mkdtemp(prefix="jpgold_")
    try:
        refio.write_restart(os.path.join(tmp, "in.rst"), s.x, s.v, lam=s.lam)
        run_md(tmp, s, dict(cmm, METHOD="CANO", RESTAR="YES", NAMERI="in.rst", NAMERO="out.rst", LOOPLI="1",
                            FDEBUG="MAKE", SETTEM="300.0"))
        dump = refio.read_force_dump(os.path.join(tmp, "1stp_force.bin"), s.n)
        out["force_grad"] = dump["grad"]
        out["force_epot"] = dump["epot"]
        t2 = os.path.join(tmp, "traj")
        os.makedirs(t2)
        refio.write_restart(os.path.join(t2, "in.rst"), s.x, s.v, lam=s.lam)
        n = opt["traj"]
        run_md(t2, s, dict(cmm, METHOD="MICR", RESTAR="YES", NAMERI="in.rst", NAMERO="out.rst", LOOPLI=str(n),
                           UPDATE="10", TIMEST="2.0", NAMETO="ene.dat", OUTLOG="100000", OUTCLE="100000"))
        r = refio.read_restart(os.path.join(t2, "out.rst"))
        out.update(traj_x=r["x"], traj_v=r["v"], traj_nsteps=np.int64(n),
                   traj_ene=np.loadtxt(os.path.join(t2, "ene.dat"), ndmin=2))
        if opt.get("mon"):
            t5 = os.path.join(tmp, "mon")
            os.makedirs(t5)
            refio.write_restart(os.path.join(t5, "in.rst"), s.x, s.v, lam=s.lam)
            log = run_md(t5, s, dict(cmm, METHOD="MICR", RESTAR="YES", NAMERI="in.rst", NAMERO="out.rst",
                                     LOOPLI=str(opt["mon"]), UPDATE="10", TIMEST="2.0", OUTLOG="20",
                                     OUTCLE="100000"))
            out["mon_micr"] = parse_monitor(log)
        if opt.get("cano"):
            t3 = os.path.join(tmp, "cano")
            os.makedirs(t3)
            refio.write_restart(os.path.join(t3, "in.rst"), s.x, s.v, lam=s.lam)
            nc = opt["cano"]
            run_md(t3, s, dict(cmm, METHOD="CANO", RESTAR="YES", NAMERI="in.rst", NAMERO="out.rst",
                               LOOPLI=str(nc), UPDATE="10", TIMEST="2.0", SETTEM="300.0", TEMPCO="NO",
                               NAMETO="ene.dat", OUTLOG="100000", OUTCLE="100000"))
            r = refio.read_restart(os.path.join(t3, "out.rst"))
            out.update(cano_x=r["x"], cano_v=r["v"], cano_nsteps=np.int64(nc),
                       cano_ene=np.loadtxt(os.path.join(t3, "ene.dat"), ndmin=2))
        if opt.get("alsd"):
            tag = opt["alsd"]
            celw, ceup, tempco, ltc, lw = ALSD_RUNS[tag]
            t4 = os.path.join(tmp, "alsd")
            os.makedirs(t4)
            refio.write_restart(os.path.join(t4, "in.rst"), s.x, s.v, lam=s.lam)
            write_alsd_params(os.path.join(t4, "ge.par"), celw, ceup)
            inp = {"LTCNTL": ltc}
            if lw is not None:
                inp["LWEIGH"] = repr(float(lw))
            run_md(t4, s, dict(cmm, METHOD="GE", RESTAR="YES", NAMERI="in.rst", NAMERO="out.rst", UPDATE="10",
                               TIMEST="2.0", SETTEM="300.0", TEMPCO=tempco, NAMEGE="ge.par", NAMETO="ene.dat",
                               LOOPLI="10", OUTLOG="100000", OUTCLE="100000"), inp=inp)
            r = refio.read_restart(os.path.join(t4, "out.rst"))
            out.update({"alsd_poly": ALSD_POLY, "alsd_low": ALSD_LOW, "alsd_high": ALSD_HIGH,
                        "alsd_walls": np.array(ALSD_WALLS), "alsd_tempce": np.array(ALSD_TEMPCE),
                        f"alsd_{tag}_x": r["x"], f"alsd_{tag}_v": r["v"], f"alsd_{tag}_lam": np.float64(r["lam"]),
                        f"alsd_{tag}_lamv": np.float64(r["lam_v"]), f"alsd_{tag}_loop": np.int64(r["loop"]),
                        f"alsd_{tag}_range": np.array([celw, ceup]),
                        f"alsd_{tag}_ene": np.loadtxt(os.path.join(t4, "ene.dat"), ndmin=2),
                        f"alsd_{tag}_opt": np.array([1 if tempco == "YES" else 0, 1 if ltc == "YES" else 0,
                                                     0.0 if lw is None else lw])})
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, **out)
    return path


# md_run's heating loop (METHOD= CANO, HEATLO= 10, STARTT= 100 K, SETTEM=
# 300 K): tag -> (restart loop count, steps, TEMPCO)
HEAT_RUNS = {"a": (0, 15, "NO"), "b": (4, 12, "YES")}


def heat_jobs(s: systems.System, tmp: str) -> dict:
    """METHOD= CANO runs through md_run's heating loop: the target
    temperature ramps from STARTT to SETTEM over HEATLO steps of md_run's
    step count, which continues from the restart's loop count."""
    out = {}
    for tag, (loop0, n, tempco) in HEAT_RUNS.items():
        d = os.path.join(tmp, "heat_" + tag)
        os.makedirs(d)
        refio.write_restart(os.path.join(d, "in.rst"), s.x, s.v, loop=loop0, lam=s.lam)
        run_md(d, s, {"METHOD": "CANO", "RESTAR": "YES", "NAMERI": "in.rst", "NAMERO": "out.rst", "LOOPLI": str(n),
                      "UPDATE": "10", "TIMEST": "2.0", "SETTEM": "300.0", "STARTT": "100.0", "HEATLO": "10",
                      "TEMPCO": tempco, "OUTLOG": "100000", "OUTCLE": "100000"})
        r = refio.read_restart(os.path.join(d, "out.rst"))
        out.update({f"heat_{tag}_x": r["x"], f"heat_{tag}_v": r["v"], f"heat_{tag}_loop": np.int64(r["loop"]),
                    f"heat_{tag}_opt": np.array([loop0, n, 1 if tempco == "YES" else 0])})
    return out


# Centre-of-mass stopping (STOPCE= TRAN/ROTA/BOTH of the first chain,
# NSTPCN= 1, remotv.f90) and md_run's initial step of a run without a
# restart (RESTAR= NO, INITIA= ZERO: v = 0, then staver's half kick):
# tag -> (METHOD, STOPCE, RESTAR, steps, TEMPCO).  md_run's remotv leaves
# the angular velocity unset under STOPCE= TRAN (remotv.f90:54-69, 122-127
# read avelcn, set only for ROTA and BOTH at 73-116), so TRAN runs of md_run
# are not fixtures.
STOP_RUNS = {"mb": ("MICR", "BOTH", "YES", 12, "NO"), "mr": ("MICR", "ROTA", "YES", 12, "NO"),
             "cb": ("CANO", "BOTH", "YES", 12, "NO"), "cy": ("CANO", "ROTA", "YES", 12, "YES"),
             "s0": ("MICR", "NO", "NO", 12, "NO"), "sc": ("CANO", "BOTH", "NO", 12, "NO")}


def stop_jobs(s: systems.System, tmp: str, extra: dict = None) -> dict:
    """md_run runs of STOP_RUNS (extra: more EXE> MD keywords, e.g. CMM)."""
    out = {}
    for tag, (method, stopce, restar, n, tempco) in STOP_RUNS.items():
        d = os.path.join(tmp, "stop_" + tag)
        os.makedirs(d)
        md = dict(extra or {}, METHOD=method, RESTAR=restar, NAMERO="out.rst", LOOPLI=str(n), UPDATE="10",
                  TIMEST="2.0", STOPCE=stopce, OUTLOG="100000", OUTCLE="100000")
        if restar == "YES":
            refio.write_restart(os.path.join(d, "in.rst"), s.x, s.v, lam=s.lam)
            md["NAMERI"] = "in.rst"
        else:
            md["INITIA"] = "ZERO"
            if s.lam:   # no restart to take lambda from (inpmd.f90:290-297)
                md["LAMBDA"] = repr(float(s.lam))
        if method == "CANO":
            md.update(SETTEM="300.0", TEMPCO=tempco)
        run_md(d, s, md)
        r = refio.read_restart(os.path.join(d, "out.rst"))
        out.update({f"stop_{tag}_x": r["x"], f"stop_{tag}_v": r["v"], f"stop_{tag}_loop": np.int64(r["loop"])})
    return out


def add_stop(name: str) -> str:
    """Adds the STOPCE and RESTAR= NO runs to an existing fixture."""
    cmm = name in CMM_RECIPES
    kind, args, opt = (CMM_RECIPES if cmm else RECIPES)[name]
    s = build_system(kind, args)
    path = os.path.join(HERE, f"{name}.npz")
    with np.load(path, allow_pickle=False) as z:
        out = {k: z[k] for k in z.files}
    assert str(out["fingerprint"]) == fingerprint(s)
    tmp = tempfile.mkdtemp(prefix="jpgold_")
    try:
        out.update(stop_jobs(s, tmp, {"CAL15M": "CMM", "BOUNDA": "NO"} if cmm else None))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    np.savez_compressed(path, **out)
    return path


def add_heat(name: str) -> str:
    """Adds the heating-loop runs to an existing fixture."""
    kind, args, opt = RECIPES[name]
    s = build_system(kind, args)
    path = os.path.join(HERE, f"{name}.npz")
    with np.load(path, allow_pickle=False) as z:
        out = {k: z[k] for k in z.files}
    assert str(out["fingerprint"]) == fingerprint(s)
    tmp = tempfile.mkdtemp(prefix="jpgold_")
    try:
        out.update(heat_jobs(s, tmp))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    np.savez_compressed(path, **out)
    return path


def add_alsd(name: str) -> str:
    """Adds the ALSD runs to an existing fixture."""
    kind, args, opt = RECIPES[name]
    s = build_system(kind, args)
    path = os.path.join(HERE, f"{name}.npz")
    with np.load(path, allow_pickle=False) as z:
        out = {k: z[k] for k in z.files}
    assert str(out["fingerprint"]) == fingerprint(s)
    tmp = tempfile.mkdtemp(prefix="jpgold_")
    try:
        out.update(alsd_jobs(s, tmp))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    np.savez_compressed(path, **out)
    return path


def add_ge(name: str) -> str:
    """Adds the METHOD= GE runs to an existing fixture."""
    kind, args, opt = RECIPES[name]
    s = build_system(kind, args)
    path = os.path.join(HERE, f"{name}.npz")
    with np.load(path, allow_pickle=False) as z:
        out = {k: z[k] for k in z.files}
    assert str(out["fingerprint"]) == fingerprint(s)
    tmp = tempfile.mkdtemp(prefix="jpgold_")
    try:
        out.update(ge_jobs(s, tmp))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    np.savez_compressed(path, **out)
    return path


def add_host(name: str) -> str:
    """Adds the md.inp host jobs' outputs to an existing fixture."""
    kind, args, opt = RECIPES[name]
    s = build_system(kind, args)
    path = os.path.join(HERE, f"{name}.npz")
    with np.load(path, allow_pickle=False) as z:
        out = {k: z[k] for k in z.files}
    assert str(out["fingerprint"]) == fingerprint(s)
    tmp = tempfile.mkdtemp(prefix="jpgold_")
    try:
        out.update(host_jobs(s, tmp))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    np.savez_compressed(path, **out)
    return path


def add_mon(name: str) -> str:
    """Adds the monitoring jobs' outputs to an existing fixture (same recipe)."""
    kind, args, opt = RECIPES[name]
    s = build_system(kind, args)
    path = os.path.join(HERE, f"{name}.npz")
    with np.load(path, allow_pickle=False) as z:
        out = {k: z[k] for k in z.files}
    assert str(out["fingerprint"]) == fingerprint(s)
    tmp = tempfile.mkdtemp(prefix="jpgold_")
    try:
        out.update(mon_jobs(s, opt, tmp))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    np.savez_compressed(path, **out)
    return path


def add_cano(name: str) -> str:
    """Adds the cano job's outputs to an existing fixture (same recipe)."""
    kind, args, opt = RECIPES[name]
    s = build_system(kind, args)
    path = os.path.join(HERE, f"{name}.npz")
    with np.load(path, allow_pickle=False) as z:
        out = {k: z[k] for k in z.files}
    assert str(out["fingerprint"]) == fingerprint(s)
    tmp = tempfile.mkdtemp(prefix="jpgold_")
    try:
        out.update(cano_job(s, opt, tmp))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    np.savez_compressed(path, **out)
    return path


def main(argv):
    if not os.path.exists(MD_RUN):
        subprocess.run(["bash", os.path.join(ROOT, "oracle", "build_ref.sh")], check=True)
    cano_only = "--cano-only" in argv
    mon_only = "--mon-only" in argv
    host_only = "--host-only" in argv
    ge_only = "--ge-only" in argv
    alsd_only = "--alsd-only" in argv
    heat_only = "--heat-only" in argv
    stop_only = "--stop-only" in argv
    cmm_all = "--cmm" in argv
    argv = [a for a in argv if a not in ("--cano-only", "--mon-only", "--host-only", "--ge-only", "--alsd-only",
                                         "--cmm", "--heat-only", "--stop-only")]
    if stop_only:
        for n in argv:
            p = add_stop(n)
            print(n, "->", os.path.relpath(p, ROOT), os.path.getsize(p), "bytes", flush=True)
        return
    names = argv or [n for n in RECIPES if not mon_only or RECIPES[n][2].get("mon")]
    if cmm_all or any(n in CMM_RECIPES for n in names):
        for n in [n for n in names if n in CMM_RECIPES] or list(CMM_RECIPES):
            p = make_cmm(n)
            print(n, "->", os.path.relpath(p, ROOT), os.path.getsize(p), "bytes", flush=True)
        return
    for n in names:
        if heat_only:
            p = add_heat(n)
        elif alsd_only:
            p = add_alsd(n)
        elif ge_only:
            p = add_ge(n)
        elif host_only:
            p = add_host(n)
        else:
            p = add_cano(n) if cano_only else (add_mon(n) if mon_only else make(n))
        print(n, "->", os.path.relpath(p, ROOT), os.path.getsize(p), "bytes", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
