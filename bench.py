"""Benchmark: ns/day and pair-interactions/s of the ZD + LJ MD step on the
100k-atom water box (33,333 waters + 60 Na/Cl pairs, 100.0 A, 12 A cutoff,
SHAKE, list rebuilt every 10 steps, dt = 2 fs) on one MI355X.

Prints one JSON line.  The timed region is K full MD steps on the device
state (pair forces, ZD excess, kick/drift, SHAKE, velocity update, list
rebuild every 10 steps, energies every 100 steps; the plain steps between
rebuilds replayed from captured HIP graphs); jp_run_nve returns after a
stream synchronize, and the host clock brackets it.  The pair kernel's
duration for the roofline comes from two HIP events around 200 force-step
pair launches back to back (jp_time_pair) that follow, on the same stream
and state.  Beside it: the plain FP64 GPU path (same protocol, fewer steps) and
the reference md_run on the host cores (bounded sample).

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
       python bench.py --workload {0,1,4}   other single-system workloads (same protocol)
       python bench.py --ensemble           64 uneven systems: one batch vs 64 sequential runs
       python bench.py --drift N            N-step NVE energy drift, FP32/fixed point vs FP64
       python bench.py --method CANO        the same runs with METHOD= CANO (velocity scaling, 300 K)
       python bench.py --cmm R              the cell multipole method (CAL15M= CMM, BOUNDA= NO) on a
                                            non-periodic water droplet of radius R A, FP64
"""
from __future__ import annotations

import argparse
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3      # MI355X FP32 vector (packed FMA counted)
PEAK_FP64_TFLOPS = 78.6       # MI355X FP64 vector
# CMM near field, algorithmic FP64 flop per ordered pair evaluated (FMA = 2,
# divide and square root 1 each): Coulomb only, Coulomb + LJ, LJ only
FLOP_CMM_EL, FLOP_CMM_LJEL, FLOP_CMM_LJ = 20, 31, 26
PEAK_HBM_GBS = 8000.0
FLOP_EL, FLOP_LJEL = 42, 54   # algorithmic flop per pair (FMA = 2, rsqrt = 1)
DT_FS = 2.0
# The synthetic chains carry light H atoms inside the backbone, held to the
# next heavy atom by a harmonic bond (omega * dt ~ 1 at 2 fs): chain-in-water
# runs use 1 fs.
DT_CHAIN_FS = 1.0
UPDATE = 10
NMON = 100


def ns_per_day(steps_per_s: float) -> float:
    return steps_per_s * DT_FS * 0.0864      # md.f90:305-306


DT_RELAX, RELAX_STEPS = 0.25, 200   # set-up stage that sheds the generated state's strain


def equilibrated(system, nsteps: int):
    """Velocity-rescaled leap-frog from the generated start (set-up only):
    RELAX_STEPS at DT_RELAX fs, then `nsteps` at the run's time step, so
    that the timed NVE segment starts from an equilibrated liquid."""
    from j_presto_amd import engine
    state = None
    for dt, n in ((DT_RELAX, RELAX_STEPS), (DT_FS, nsteps)):
        md = engine.MdSystem(system, dt=dt, precision=engine.FP32)
        if state is not None:
            md.set_state(*state)
        ier = md.celprp()
        assert ier == 0, ier
        ier = engine.equilibrate(md, n)
        assert ier == 0, (dt, ier)
        state = md.get_state()
        md.close()
    return state


def run_gpu(system, precision, steps, warmup, state=None, pair_steps=200, autotune=True, cano=False):
    from j_presto_amd import engine, native
    md = engine.MdSystem(system, dt=DT_FS, precision=precision)
    if state is not None:
        md.set_state(*state)
    if cano:   # METHOD= CANO, SETTEM= 300 K, TEMPCO= NO
        md.set_canonical(300.0)
    ier = md.celprp()
    assert ier == 0, ier
    if autotune:   # launch shapes timed by a zero-step run, before the warm-up
        md.set_option(native.JP_OPT_AUTOTUNE, 1)
        _, ier = md.runmd(0, update=UPDATE)
        assert ier == 0, ier
    if warmup:
        _, ier = md.runmd(warmup, update=UPDATE)
        assert ier == 0, ier
    md.celprp()
    # the timed run starts at a step count that is a multiple of the list and
    # energy intervals, and every step graph it replays is captured beforehand
    md.set_step(0)
    md.prepare_graphs(steps, update=UPDATE, nmon=NMON)
    t0 = time.perf_counter()
    log, ier = md.runmd(steps, update=UPDATE, nmon=NMON)
    t1 = time.perf_counter()
    assert ier == 0, ier
    captures = md.get_option(native.JP_OPT_GRAPH_CAPTURES)
    tm = md.timing()
    n, nlj, nlisted = md.count_pairs(detail=True)
    # pair-kernel duration: HIP events around pair_steps launches back to back
    tp = md.time_pair(pair_steps)
    shapes = {"pair_waves": md.get_option(native.JP_OPT_PAIR_WAVES),
              "list_waves": md.get_option(native.JP_OPT_LIST_WAVES),
              "list_batches": md.get_option(native.JP_OPT_LIST_BATCH),
              "grid_shift": md.get_option(native.JP_OPT_GRID_SHIFT), "autotuned": bool(autotune)}
    md.close()
    return dict(wall=t1 - t0, timing=tm, pairs=n, pairs_lj=nlj, listed=nlisted, log=log, t_pair=tp, shapes=shapes,
                graph_captures_timed=captures)


def accuracy(system, state):
    """FP32/fixed-point forces and energies against the plain FP64 GPU path
    (itself within 1e-12 of the FP64 oracle) at the benchmark's start state:
    relative RMS of the 3N force components, relative ELE15 and VDW15 error."""
    import numpy as np
    from j_presto_amd import engine
    out = {}
    res = {}
    for name, prec in (("fp64", engine.FP64), ("fp32", engine.FP32)):
        md = engine.MdSystem(system, dt=DT_FS, precision=prec)
        md.set_state(*state)
        assert md.celprp() == 0
        g, et, ier = md.bonded()
        assert ier == 0
        g, ee, ev, eh, ier = md.cmmdo(grad=g)
        assert ier == 0
        md.close()
        res[name] = (g, ee.sum(), ev.sum())
    g64, e64, v64 = res["fp64"]
    g32, e32, v32 = res["fp32"]
    out["force_rel_rms_fp32_vs_fp64"] = float(np.sqrt(((g32 - g64) ** 2).sum() / (g64 ** 2).sum()))
    out["ele15_rel_err"] = float(abs(e32 - e64) / abs(e64))
    out["vdw15_rel_err"] = float(abs(v32 - v64) / abs(v64))
    out["ele15_fp64"] = float(e64)
    out["vdw15_fp64"] = float(v64)
    return out


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_reference(system, nsteps: int):
    """The reference md_run (oracle/_ref/md_run, built from j_presto's own
    sources) on this host: PARALV= DOUB, N_SIMD= 64, all cores."""
    exe = os.path.join(ROOT, "oracle", "_ref", "md_run")
    if not os.path.exists(exe):
        return None
    from j_presto_amd import refio
    cores = len(os.sched_getaffinity(0))
    tmp = tempfile.mkdtemp(prefix="jpcpu_")
    try:
        files = refio.write_inputs(tmp, system)
        refio.write_restart(os.path.join(tmp, "in.rst"), system.x, system.v)
        refio.write_md_inp(os.path.join(tmp, "md.inp"), files, system, {
            "METHOD": "MICR", "RESTAR": "YES", "NAMERI": "in.rst", "NAMERO": "out.rst",
            "LOOPLI": str(nsteps), "UPDATE": str(UPDATE), "TIMEST": str(DT_FS), "OUTLOG": "100000",
            "OUTCLE": "100000", "N_SIMD": "64"})
        env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_STACKSIZE="512M")
        with open(os.path.join(tmp, "md.inp")) as fin, open(os.path.join(tmp, "md.log"), "w") as fout:
            subprocess.run(["bash", "-c", f"ulimit -s unlimited; exec {exe}"], stdin=fin, stdout=fout,
                           stderr=subprocess.STDOUT, cwd=tmp, env=env, check=True, timeout=900)
        log = open(os.path.join(tmp, "md.log")).read()
        speed = float(re.search(r"SPEED \(ns/day\)\s*:\s*([0-9.Ee+-]+)", log).group(1))
        loop = float(re.search(r"^\s+MD LOOP\s+:\s+([0-9.Ee+-]+)", log, re.M).group(1))
        return dict(value=speed, unit="ns/day", cores=cores, kind="reference", cpu_model=cpu_model(),
                    sample=f"md_run (flang -O2 build of the reference sources), {nsteps} MICR steps of this "
                           f"workload, PARALV=DOUB N_SIMD=64, {cores} OpenMP threads, MD loop {loop:.2f} s")
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def traffic_from_profile():
    """HBM bytes per pair-kernel launch from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated), if present."""
    p = os.path.join(ROOT, "profiles", "pair_kernel_water100k_traffic.json")
    if os.path.exists(p):
        d = json.load(open(p))
        return d.get("bytes_per_launch")
    return None


def calibrate_copy(nbytes: int = 1 << 30, reps: int = 5):
    """Plain device-to-device copies of a known size: the reference point
    for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes."""
    import torch
    src = torch.empty(nbytes, dtype=torch.uint8, device="cuda").fill_(1)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": nbytes, "reps": reps}), flush=True)


def drift_fit(log, natom):
    """Least-squares slope of E_tot(t) in kcal/mol/ns, per 1,000 atoms too,
    and the span and RMS fluctuation about the line."""
    import numpy as np
    t_ns = log["step"] * DT_FS * 1e-6
    e = log["etot"]
    a, b = np.polyfit(t_ns, e, 1)
    res = e - (a * t_ns + b)
    return {"slope_kcal_mol_ns": float(a), "slope_per_1000_atoms": float(a * 1000.0 / natom),
            "span_kcal_mol": float(e.max() - e.min()), "rms_fluct_kcal_mol": float(res.std()),
            "etot_first": float(e[0]), "etot_last": float(e[-1]), "samples": int(len(e))}


def cmm_pair_counts(s, cells, idelay: int, rlim: float):
    """Ordered pairs the CMM near-field kernels evaluate at the rebuild:
    Coulomb-only and Coulomb + LJ pairs of the 27-cell neighbourhoods
    (excluded pairs and i = j left out), LJ pairs of the second-nearest
    cells within rlim."""
    import numpy as np
    from scipy.spatial import cKDTree
    vdw = np.array([t.eps != 0.0 for t in s.types])[s.atype - 1]
    nv = int(cells.max()) + 2
    c = cells - 1
    lin = (c[:, 2] * nv + c[:, 1]) * nv + c[:, 0]
    cnt = np.bincount(lin, minlength=nv ** 3).reshape(nv, nv, nv)
    cntv = np.bincount(lin[vdw], minlength=nv ** 3).reshape(nv, nv, nv)

    def box27(a):
        p = np.pad(a, 1)
        out = np.zeros_like(a)
        for dz in range(3):
            for dy in range(3):
                for dx in range(3):
                    out += p[dz:dz + nv, dy:dy + nv, dx:dx + nv]
        return out
    n27, n27v = box27(cnt), box27(cntv)
    tot = n27[c[:, 2], c[:, 1], c[:, 0]] - 1
    ljel = np.where(vdw, n27v[c[:, 2], c[:, 1], c[:, 0]] - 1, 0)
    ex = s.excl14()
    cheb = np.abs(c[ex[:, 0]] - c[ex[:, 1]]).max(1)
    near_ex = int(2 * (cheb <= 1).sum())
    near_ex_v = int(2 * ((cheb <= 1) & vdw[ex[:, 0]] & vdw[ex[:, 1]]).sum())
    n_all = int(tot.sum()) - near_ex
    n_ljel = int(ljel.sum()) - near_ex_v
    iv = np.where(vdw)[0]
    pairs = cKDTree(s.x[iv]).query_pairs(rlim, output_type="ndarray")
    a, b = iv[pairs[:, 0]], iv[pairs[:, 1]]
    d = s.x[a] - s.x[b]
    ch = np.abs(c[a] - c[b]).max(1)
    lj2 = ((ch >= 2) & (ch <= idelay) & ((d * d).sum(1) < rlim * rlim))
    return {"el": n_all - n_ljel, "ljel": n_ljel, "lj2": int(2 * lj2.sum())}


def cmm_cpu_reference(system, nsteps: int):
    """md_run (oracle/_ref/md_run) on the droplet with CAL15M= CMM, BOUNDA= NO
    on all host cores: a bounded sample of MICR steps."""
    exe = os.path.join(ROOT, "oracle", "_ref", "md_run")
    if not os.path.exists(exe):
        return None
    from j_presto_amd import refio
    cores = len(os.sched_getaffinity(0))
    tmp = tempfile.mkdtemp(prefix="jpcmm_")
    try:
        files = refio.write_inputs(tmp, system)
        refio.write_restart(os.path.join(tmp, "in.rst"), system.x, system.v)
        refio.write_md_inp(os.path.join(tmp, "md.inp"), files, system, {
            "METHOD": "MICR", "RESTAR": "YES", "NAMERI": "in.rst", "NAMERO": "out.rst", "CAL15M": "CMM",
            "BOUNDA": "NO", "LOOPLI": str(nsteps), "UPDATE": str(UPDATE), "TIMEST": str(DT_FS),
            "OUTLOG": "100000", "OUTCLE": "100000"})
        env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_STACKSIZE="512M")
        with open(os.path.join(tmp, "md.inp")) as fin, open(os.path.join(tmp, "md.log"), "w") as fout:
            subprocess.run(["bash", "-c", f"ulimit -s unlimited; exec {exe}"], stdin=fin, stdout=fout,
                           stderr=subprocess.STDOUT, cwd=tmp, env=env, check=True, timeout=900)
        log = open(os.path.join(tmp, "md.log")).read()
        speed = float(re.search(r"SPEED \(ns/day\)\s*:\s*([0-9.Ee+-]+)", log).group(1))
        loop = float(re.search(r"^\s+MD LOOP\s+:\s+([0-9.Ee+-]+)", log, re.M).group(1))
        return dict(value=speed, unit="ns/day", cores=cores, kind="reference", cpu_model=cpu_model(),
                    sample=f"md_run (flang -O2 build of the reference sources), {nsteps} MICR steps of this "
                           f"droplet with CAL15M=CMM BOUNDA=NO, {cores} OpenMP threads, MD loop {loop:.2f} s")
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def run_cmm(R: float, steps: int, warmup: int, equil: int, cpu_steps: int, no_cpu: bool):
    """CAL15M= CMM (BOUNDA= NO) MICR steps on a water droplet of radius R:
    relaxed and equilibrated with the CMM forces themselves, then K timed
    steps (far field every CMMUPD= 5 steps, tables every 10, energies every
    100), FP64 as md_run."""
    import numpy as np
    from j_presto_amd import engine, systems
    s = systems.water_droplet(R, seed=97)
    cminsiz, nlev, cmmupd = 6.0, 3, 5
    state = None
    for dt, n in ((DT_RELAX, RELAX_STEPS), (DT_FS, equil)):
        md = engine.MdSystem(s, dt=dt, precision=engine.FP64)
        if state is not None:
            md.set_state(*state)
        md.set_cmm(cminsiz, nlev, cmmupd)
        assert md.celprp() == 0
        assert engine.equilibrate(md, n) == 0
        state = md.get_state()
        md.close()
    md = engine.MdSystem(s, dt=DT_FS, precision=engine.FP64)
    md.set_state(*state)
    md.set_cmm(cminsiz, nlev, cmmupd)
    assert md.celprp() == 0
    if warmup:
        _, ier = md.runmd(warmup, update=UPDATE)
        assert ier == 0, ier
    assert md.celprp() == 0
    md.set_step(0)
    md.prepare_graphs(steps, update=UPDATE, nmon=NMON)
    t0 = time.perf_counter()
    log, ier = md.runmd(steps, update=UPDATE, nmon=NMON)
    t1 = time.perf_counter()
    assert ier == 0, ier
    tm = md.timing()
    assert md.celprp() == 0
    t_near, t_far = md.time_cmm(50)
    cells = md.cmm_cells()
    x_end, v_end = md.get_state()
    md.close()
    s.x = x_end
    nv1 = int(cells.max())
    idelay = int(np.ceil(s.cutoff / cminsiz))
    pc = cmm_pair_counts(s, cells, idelay, s.cutoff)
    flop = pc["el"] * FLOP_CMM_EL + pc["ljel"] * FLOP_CMM_LJEL + pc["lj2"] * FLOP_CMM_LJ
    achieved = flop / t_near / 1e12
    sps = steps / (t1 - t0)
    et = log["etot"]
    out = {
        "metric": f"ns/day, CMM (CAL15M=CMM, BOUNDA=NO) water droplet R={R:g} A, FP64",
        "value": round(ns_per_day(sps), 3), "unit": "ns/day", "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": round(1e3 * (t1 - t0) / steps, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"water droplet R={R:g} A ({s.n} atoms, non-periodic)", "natom": s.n,
                   "cutoff_A": s.cutoff, "celsiz_A": cminsiz, "cellvl": nlev, "cmmupd": cmmupd,
                   "cells_per_edge": nv1, "dt_fs": DT_FS, "list_update": UPDATE, "energy_every": NMON,
                   "method": "MICR"},
        "pairs_per_step": pc,
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64_TFLOPS, 5), "traffic": None,
                     "kernel": "k_cmm_near + k_cmm_lj2", "avg_launch_us": round(t_near * 1e6, 3),
                     "flop_per_launch": flop},
        "far_field_us": round(t_far * 1e6, 3),
        "step_breakdown_ms": {"near": round(1e3 * t_near, 4), "far_per_step": round(1e3 * t_far / cmmupd, 4),
                              "rebuild_per_step": round(1e3 * tm["list"] / steps, 4),
                              "rest": round(1e3 * (t1 - t0) / steps - 1e3 * (t_near + t_far / cmmupd)
                                            - 1e3 * tm["list"] / steps, 4)},
        "nve_etot_span_kcal_mol": float(et.max() - et.min()) if len(et) else None,
        "cpu_baseline": None,
    }
    if not no_cpu:
        s.x, s.v = state
        out["cpu_baseline"] = cmm_cpu_reference(s, cpu_steps)
    print(json.dumps(out), flush=True)


def run_drift(nsteps: int, equil: int, reps: int = 1):
    """N-step NVE runs in both precisions from the same equilibrated state;
    with reps > 1, further start states 1,000 rescaled steps apart, and the
    mean and spread of the slopes."""
    import numpy as np
    from j_presto_amd import engine, systems
    s = systems.workload(2)
    state = equilibrated(s, equil)
    out = {"mode": "nve-drift", "workload": systems.WORKLOADS[2], "natom": s.n, "steps": nsteps,
           "dt_fs": DT_FS, "energy_every": NMON, "list_update": UPDATE, "data": "synthetic", "reps": reps}
    slopes = {"fp32_fixed": [], "fp64": []}
    for rep in range(reps):
        if rep:
            md = engine.MdSystem(s, dt=DT_FS, precision=engine.FP32)
            md.set_state(*state)
            assert md.celprp() == 0
            assert engine.equilibrate(md, 1000) == 0
            state = md.get_state()
            md.close()
        for name, prec in (("fp32_fixed", engine.FP32), ("fp64", engine.FP64)):
            md = engine.MdSystem(s, dt=DT_FS, precision=prec)
            md.set_state(*state)
            assert md.celprp() == 0
            t0 = time.perf_counter()
            log, ier = md.runmd(nsteps, update=UPDATE, nmon=NMON)
            wall = time.perf_counter() - t0
            assert ier == 0, ier
            md.close()
            fit = drift_fit(log, s.n)
            fit["ns_per_day"] = round(ns_per_day(nsteps / wall), 3)
            slopes[name].append(fit["slope_per_1000_atoms"])
            if rep == 0:
                out[name] = fit
            print(f"# rep {rep} {name}: {fit}", file=sys.stderr, flush=True)
    if reps > 1:
        out["slope_per_1000_atoms"] = {k: {"mean": float(np.mean(v)), "std": float(np.std(v, ddof=1)),
                                           "all": [round(x, 4) for x in v]} for k, v in slopes.items()}
    print(json.dumps(out), flush=True)


def run_ensemble(steps: int, warmup: int, equil: int):
    """64 uneven chain-in-water systems (3k-30k atoms, each with its own
    lambda): one batch sharing every launch against 64 separate runs, both
    from the same equilibrated states."""
    from j_presto_amd import engine, systems
    syss = systems.ensemble(64)
    natom = sum(s.n for s in syss)
    states = None
    for dt, n in ((DT_RELAX, RELAX_STEPS), (DT_FS, equil)):
        ens = engine.Ensemble(syss, dt=dt, precision=engine.FP32)
        for i, st in enumerate(states or []):
            ens.set_state(i, *st)
        ier = engine.equilibrate_ensemble(ens, n)
        assert ier == 0, (dt, ier)
        states = [ens.get_state(i) for i in range(len(syss))]
        if dt != DT_FS:
            ens.close()
    from j_presto_amd import native
    ens.set_option(native.JP_OPT_AUTOTUNE, 1)
    _, ier = ens.runmd(0, update=UPDATE)
    assert ier == 0, ier
    shapes = {"pair_waves": ens.get_option(native.JP_OPT_PAIR_WAVES),
              "list_waves": ens.get_option(native.JP_OPT_LIST_WAVES),
              "list_batches": ens.get_option(native.JP_OPT_LIST_BATCH),
              "grid_shift": ens.get_option(native.JP_OPT_GRID_SHIFT), "autotuned": True}
    _, ier = ens.runmd(warmup, update=UPDATE)
    assert ier == 0, ier
    t0 = time.perf_counter()
    _, ier = ens.runmd(steps, update=UPDATE)
    wb = time.perf_counter() - t0
    assert ier == 0, ier
    ens.close()
    ws = 0.0
    for s, st in zip(syss, states):
        md = engine.MdSystem(s, dt=DT_FS, precision=engine.FP32)
        md.set_state(*st)
        assert md.celprp() == 0
        md.set_option(native.JP_OPT_AUTOTUNE, 1)   # each sequential run tuned for its own size
        _, ier = md.runmd(0, update=UPDATE)
        assert ier == 0, ier
        _, ier = md.runmd(warmup, update=UPDATE)
        assert ier == 0, ier
        t0 = time.perf_counter()
        _, ier = md.runmd(steps, update=UPDATE)
        ws += time.perf_counter() - t0
        assert ier == 0, ier
        md.close()
    agg_b = ns_per_day(len(syss) * steps / wb)
    agg_s = ns_per_day(len(syss) * steps / ws)
    print(json.dumps({
        "metric": "aggregate ns/day, 64-system ensemble", "value": round(agg_b, 3), "unit": "ns/day",
        "n_gpus": 1, "steps": steps, "warmup": warmup, "higher_is_better": True, "dtype": "fp32",
        "data": "synthetic", "config": {"workload": systems.WORKLOADS[3], "nsys": len(syss), "natom_total": natom,
                                        "sizes_min_max": [min(s.n for s in syss), max(s.n for s in syss)],
                                        "precision": "FP32 pair math, int64 fixed-point accumulation"},
        "batched_wall_s": round(wb, 4), "sequential_wall_s": round(ws, 4),
        "sequential_ns_per_day": round(agg_s, 3), "batch_speedup": round(ws / wb, 3), "launch_shapes": shapes,
        "dt_fs": DT_FS, "equilibration_steps": equil}), flush=True)


def main():
    global DT_FS
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", type=int, default=2, choices=[0, 1, 2, 4])
    ap.add_argument("--ensemble", action="store_true")
    ap.add_argument("--cmm", type=float, default=0.0, help="CMM water droplet of this radius (A) instead")
    ap.add_argument("--drift", type=int, default=0, help="run an N-step NVE drift comparison instead")
    ap.add_argument("--drift-reps", type=int, default=1, help="independent start states for --drift")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run 1 GiB device copies (profiler traffic calibration)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fp64-steps", type=int, default=0)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--equil", type=int, default=1000, help="velocity-rescaled set-up steps before timing")
    ap.add_argument("--method", default="MICR", choices=["MICR", "CANO"],
                    help="md_run METHOD of the timed runs (CANO: 300 K, TEMPCO= NO)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)
    if args.ensemble or args.workload == 1:
        DT_FS = DT_CHAIN_FS
    if args.drift:
        run_drift(args.drift, args.equil, args.drift_reps)
        return
    if args.cmm > 0:
        run_cmm(args.cmm, args.steps, args.warmup, args.equil, args.cpu_steps, args.no_cpu)
        return
    if args.ensemble:
        run_ensemble(args.steps, args.warmup, args.equil)
        return
    rank = int(os.environ.get("RANK", "0"))
    from j_presto_amd import engine, systems
    s = systems.workload(args.workload)
    state = equilibrated(s, args.equil)
    cano = args.method == "CANO"
    r32 = run_gpu(s, engine.FP32, args.steps, args.warmup, state, cano=cano)
    n64 = args.fp64_steps or max(args.steps // 4, 100)
    r64 = run_gpu(s, engine.FP64, n64, min(args.warmup, 50), state, cano=cano)
    sps = args.steps / r32["wall"]
    value = ns_per_day(sps)
    v64 = ns_per_day(n64 / r64["wall"])
    # pair kernel roofline
    n_el = r32["pairs"] - r32["pairs_lj"]
    flop = n_el * FLOP_EL + r32["pairs_lj"] * FLOP_LJEL
    t_pair = r32["t_pair"]
    achieved = flop / t_pair / 1e12
    # list build: 64 B per atom + 4 B per unique listed pair, per rebuild
    nreb = max(1, args.steps // UPDATE)
    t_list = r32["timing"]["list"] / nreb
    list_bytes = 64 * s.n + 4 * r32["listed"]
    et = r32["log"]["etot"]
    out = {
        "metric": ("ns/day + pair-interactions/s, 100k-atom ZD box; FP32/fixed-pt vs FP64 GPU" if args.workload == 2
                   else f"ns/day + pair-interactions/s, {systems.WORKLOADS[args.workload]}")
                  + ("; METHOD= CANO" if cano else ""),
        "value": round(value, 3),
        "unit": "ns/day",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * r32["wall"] / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": systems.WORKLOADS[args.workload], "natom": s.n, "cutoff_A": 12.0, "dt_fs": DT_FS,
                   "list_update": UPDATE, "energy_every": NMON, "precision": "FP32 pair math, int64 fixed-point "
                   "accumulation, FP64 SHAKE/integration", "launch_shapes": r32["shapes"],
                   "method": args.method},
        "pair_interactions_per_s": round(r32["pairs"] * sps, 1),
        "pairs_within_cutoff": r32["pairs"],
        "fp64_gpu_ns_per_day": round(v64, 3),
        "speedup_vs_fp64_gpu": round(value / v64, 3),
        "nve_etot_span_kcal_mol": float(et.max() - et.min()) if len(et) else None,
        "nve_etot_slope_kcal_mol_per_ns": (float(__import__("numpy").polyfit(r32["log"]["step"] * DT_FS * 1e-6, et, 1)[0])
                                           if len(et) > 2 else None),
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 5),
                     "traffic": traffic_from_profile() if args.workload == 2 else None,
                     "kernel": f"k_pair_f32<false, false, false, {r32['shapes']['pair_waves']}>",
                     "avg_launch_us": round(t_pair * 1e6, 3),
                     "flop_per_launch": flop},
        "list_roofline": {"bound": "hbm", "achieved": round(list_bytes / t_list / 1e9, 2), "peak": PEAK_HBM_GBS,
                          "unit": "GB/s", "frac": round(list_bytes / t_list / 1e9 / PEAK_HBM_GBS, 5),
                          "bytes_per_rebuild": list_bytes, "avg_rebuild_us": round(t_list * 1e6, 2)},
        "step_breakdown_ms": {"pair": round(1e3 * t_pair, 4), "list_per_step": round(1e3 * r32["timing"]["list"] / args.steps, 4),
                              "rest": round(1e3 * (r32["wall"] - r32["timing"]["list"]) / args.steps - 1e3 * t_pair, 4)},
        "graph_captures_timed": r32["graph_captures_timed"],
        "accuracy_vs_fp64": accuracy(s, state),
        "cpu_baseline": None,
    }
    if rank == 0 and not args.no_cpu and args.workload == 2 and not cano:
        cb = cpu_reference(s, args.cpu_steps)
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
