"""Host mirror of md_run's call surface for the ZD path, over libjpgpu.

Same routine names, argument meaning and error behaviour as the reference
(ier codes are returned, not raised):

  MdSystem.celprp()          celprp.f90:2      (bin + pair list; ier 0/28/51/99)
  MdSystem.cmmdo()           cmmdo.f90:2       (ZD+LJ pair terms, self, excess)
  MdSystem.dnergy()          dnergy.f90:2      (nonbonded energy surface:
                                                 eneval(15), enevcc(15,nfrg), grad)
  MdSystem.shake(r1, r2)     shake.f90:2       (ier 0/1/-2)
  MdSystem.runmd(n, ...)     md.f90:472 + verlet.f90:2 (METHOD= MICR)
  Ensemble                   j_presto_mpi:27-44 (N independent runs) as one
                             batched launch

Energy slots follow main.f90:157-159: 1 TOTAL, 8 VDW15, 9 ELE15, 10 HYD.
Runtime (HIP) failures raise native.JpError.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence

import numpy as np

from . import native
from .abi import DescHolder, EnergyLog, SysDesc
from .systems import System

FP64 = 0
FP32 = 1
MAXENE = 15


def _f64(a):
    return a.ctypes.data_as(native.f64p) if a is not None else None


def _i32(a):
    return a.ctypes.data_as(native.i32p) if a is not None else None



# md_run's STOPCE values (inpmd.f90:124 idfstp; NO is off)
STOPCE = {"NO": 0, "TRAN": 1, "ROTA": 2, "BOTH": 3}

class MdSystem:
    """One system on the GPU (a jp_ctx)."""

    def __init__(self, system: System, dt: float = 2.0, precision: int = FP64, **desc_kw):
        self.lib = native.load()
        self.system = system
        self.h = DescHolder(system, dt=dt, **desc_kw)
        self.n = self.h.n
        self.nfrg = self.h.nfrg
        self.precision = precision
        p = C.c_void_p()
        native.check("jp_ctx_create", self.lib.jp_ctx_create(C.byref(self.h.desc), C.byref(p)))
        self.ctx = p
        v = np.ascontiguousarray(system.v, np.float64)
        self.set_state(None, v)

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.jp_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state ---------------------------------------------------------
    def set_state(self, cord=None, vel=None):
        cord = None if cord is None else np.ascontiguousarray(cord, np.float64)
        vel = None if vel is None else np.ascontiguousarray(vel, np.float64)
        native.check("jp_set_state", self.lib.jp_set_state(self.ctx, _f64(cord), _f64(vel)))

    def get_state(self):
        x = np.zeros((self.n, 3))
        v = np.zeros((self.n, 3))
        native.check("jp_get_state", self.lib.jp_get_state(self.ctx, _f64(x), _f64(v)))
        return x, v

    # -- reference routines --------------------------------------------
    def celprp(self, cord=None) -> int:
        cord = None if cord is None else np.ascontiguousarray(cord, np.float64)
        return native.check("jp_build_list", self.lib.jp_build_list(self.ctx, _f64(cord)))

    def get_list(self):
        """(cell (N,3) md_run numbering, ptr (N+1), idx 1-based ascending)."""
        n = C.c_int64()
        native.check("jp_get_list", self.lib.jp_get_list(self.ctx, None, C.byref(n), None, None, 0))
        cell = np.zeros((self.n, 3), np.int32)
        ptr = np.zeros(self.n + 1, np.int32)
        idx = np.zeros(max(n.value, 1), np.int32)
        native.check("jp_get_list", self.lib.jp_get_list(
            self.ctx, cell.ctypes.data_as(native.i32p), C.byref(n), ptr.ctypes.data_as(native.i32p),
            idx.ctypes.data_as(native.i32p), len(idx)))
        return cell, ptr, idx[:n.value]

    def cmm_cells(self):
        """Under CMM: the level-1 cell (1-based per axis) of every atom at the
        last rebuild, (N, 3)."""
        n = C.c_int64()
        cell = np.zeros((self.n, 3), np.int32)
        native.check("jp_get_list", self.lib.jp_get_list(self.ctx, cell.ctypes.data_as(native.i32p), C.byref(n),
                                                          None, None, 0))
        return cell

    def cmmdo(self, cord=None, grad=None, precision: Optional[int] = None):
        """Returns (grad(nfrg,N,3) accumulated, eelecn, evancn, ehydcn, ier)."""
        prec = self.precision if precision is None else precision
        cord = None if cord is None else np.ascontiguousarray(cord, np.float64)
        g = np.zeros((self.nfrg, self.n, 3)) if grad is None else grad
        ee, ev, eh = np.zeros(self.nfrg), np.zeros(self.nfrg), np.zeros(self.nfrg)
        ier = native.check("jp_nonbond", self.lib.jp_nonbond(self.ctx, _f64(cord), _f64(g), _f64(ee), _f64(ev),
                                                              _f64(eh), prec))
        return g, ee, ev, eh, ier

    def bonded(self, cord=None, grad=None):
        """dnergv_PB's bonded part: (grad(nfrg,N,3) accumulated, eterm(nfrg,6)
        with columns bond, angle, torsion, improper, 1-4 vdW, 1-4 elec, ier)."""
        cord = None if cord is None else np.ascontiguousarray(cord, np.float64)
        g = np.zeros((self.nfrg, self.n, 3)) if grad is None else grad
        et = np.zeros((self.nfrg, 6))
        ier = native.check("jp_bonded", self.lib.jp_bonded(self.ctx, _f64(cord), _f64(g), _f64(et)))
        return g, et, ier

    def dnergy(self, cord=None, precision: Optional[int] = None):
        """dnergy on this path: (eneval(15), enevcc(15,nfrg), grad, ier) with
        slots 2-7 BOND, ANGLE, TORSION, IMPROPER, 1-4 VDW, 1-4 ELEC, 8 VDW15,
        9 ELE15, 10 HYD and slot 1 their sum (dnergy.f90:45-79: grad zeroed,
        the bonded terms, then cmmdo)."""
        g, et, ier = self.bonded(cord)
        if ier:
            return None, None, g, ier
        g, ee, ev, eh, ier = self.cmmdo(None, grad=g, precision=precision)
        enevcc = np.zeros((MAXENE, self.nfrg))
        enevcc[1:7] = et.T
        enevcc[8] = ee
        enevcc[7] = ev
        enevcc[9] = eh
        enevcc[0] = enevcc[1:].sum(0)
        return enevcc.sum(1), enevcc, g, ier

    def shake(self, r1, r2) -> int:
        """Constrain r1 (x(t+dt)) in place against r2 (x(t))."""
        assert r1.flags.c_contiguous and r1.dtype == np.float64
        r2 = np.ascontiguousarray(r2, np.float64)
        return native.check("jp_shake", self.lib.jp_shake(self.ctx, _f64(r1), _f64(r2)))

    def runmd(self, nsteps: int, update: int = 10, nmon: int = 0, precision: Optional[int] = None):
        """md_run's MD loop on the device state (METHOD= MICR, or CANO after
        set_canonical).  Returns (log, ier)."""
        prec = self.precision if precision is None else precision
        nl = nsteps // nmon + 1 if nmon > 0 else 0
        log = (EnergyLog * max(nl, 1))()
        ier = native.check("jp_run_nve", self.lib.jp_run_nve(self.ctx, nsteps, update, nmon, prec, log))
        out = log_to_array(log, nl)
        return out[out["step"] >= 0], ier

    def prepare_graphs(self, nsteps: int, update: int = 10, nmon: int = 0, precision: Optional[int] = None):
        """Capture the step graphs runmd(nsteps, update, nmon) will replay."""
        prec = self.precision if precision is None else precision
        native.check("jp_prepare_graphs", self.lib.jp_prepare_graphs(self.ctx, nsteps, update, nmon, prec))

    def restart_step(self, precision: Optional[int] = None):
        """md_run's restart set-up (inirst, md.f90:1104-1263): one step whose
        energies are returned as a one-record log, then the coordinates back
        and the velocities those of the step's last substep start.  Returns
        (log, ier)."""
        prec = self.precision if precision is None else precision
        log = (EnergyLog * 1)()
        ier = native.check("jp_restart_step", self.lib.jp_restart_step(self.ctx, prec, log))
        return log_to_array(log, 1), ier

    def set_step(self, idloop: int):
        """md_run's step counter (a restart's loop number); list rebuilds and
        monitoring steps follow it."""
        native.check("jp_set_step", self.lib.jp_set_step(self.ctx, int(idloop)))

    def get_step(self) -> int:
        v = C.c_int64()
        native.check("jp_get_step", self.lib.jp_get_step(self.ctx, C.byref(v)))
        return v.value

    def set_graphs(self, on: bool):
        """Plain steps of runmd as captured HIP graphs (default) or launched
        one kernel at a time; the results are the same."""
        native.check("jp_set_option", self.lib.jp_set_option(self.ctx, native.JP_OPT_GRAPHS, int(bool(on))))

    def set_option(self, option: int, value: int):
        native.check("jp_set_option", self.lib.jp_set_option(self.ctx, option, int(value)))

    def set_mcmd(self, ge: Optional[dict]):
        """METHOD= GE (multicanonical) on top of set_canonical: ge holds c
        (nwindow, ndeg+1), low, high (nwindow), alphalw, alphaup, celw, ceup
        and tempce (the NAMEGE= file); None switches it off."""
        if ge is None:
            native.check("jp_set_mcmd", self.lib.jp_set_mcmd(self.ctx, 0, 0, 0, None, None, None,
                                                             0.0, 0.0, 0.0, 0.0, 0.0))
            return
        c = np.ascontiguousarray(ge["c"], np.float64)
        low = np.ascontiguousarray(ge["low"], np.float64)
        high = np.ascontiguousarray(ge["high"], np.float64)
        self._ge = (c, low, high)
        native.check("jp_set_mcmd", self.lib.jp_set_mcmd(
            self.ctx, 1, c.shape[0], c.shape[1] - 1, _f64(c), _f64(low), _f64(high), float(ge["alphalw"]),
            float(ge["alphaup"]), float(ge["celw"]), float(ge["ceup"]), float(ge["tempce"])))

    def set_alsd(self, al: Optional[dict]):
        """ALSD (METHOD= GE under GEPSME= LMD) on top of set_canonical: al
        holds c (nwindow, ndeg+1), low, high (nwindow), lower, upper, celw,
        ceup, tempce (the NAMEGE= file), lmass (LWEIGH; <= 0: md_run's
        default) and tctl (LTCNTL); None switches it off."""
        self._al = _alsd_call(self.lib.jp_set_alsd, "jp_set_alsd", self.ctx, al)

    def set_heating(self, heatlo: int, startt: float = 300.0):
        """md_run's heating loop (HEATLO steps from STARTT K to SETTEM) for
        the canonical runs that follow; heatlo = 0 switches it off."""
        native.check("jp_set_heating", self.lib.jp_set_heating(self.ctx, int(heatlo), float(startt)))

    def set_stopce(self, mode: str, nstpcn: int = 1):
        """STOPCE= TRAN, ROTA, BOTH or NO over the first nstpcn chains
        (molecule copies) for the steps that follow; before set_canonical."""
        s = self.system
        native.check("jp_set_stopce", self.lib.jp_set_stopce(
            self.ctx, STOPCE[mode], s.chain_end(nstpcn) if mode != "NO" else 0, s.first_kind_atoms()))

    def start_step(self, precision: Optional[int] = None) -> int:
        """md_run's initial step of a run without a restart (staver) from the
        context's x and v; returns ier."""
        prec = self.precision if precision is None else precision
        return native.check("jp_start_step", self.lib.jp_start_step(self.ctx, prec))

    def set_cmm(self, cminsiz: Optional[float] = 6.0, nlev: int = 3, cmmupd: int = 5):
        """CAL15M= CMM (cell multipole method, BOUNDA= NO) for the following
        evaluations and runs: CELSIZ, CELLVL and CMMUPD; cminsiz=None returns
        to ZD.  A rebuild (celprp) follows before the next forces."""
        on = 0 if cminsiz is None else 1
        native.check("jp_set_cmm", self.lib.jp_set_cmm(self.ctx, on, float(cminsiz or 0.0), int(nlev), int(cmmupd)))

    def time_cmm(self, reps: int = 50):
        """Under CMM: (near-field, far-field) kernel seconds per launch."""
        a, b = C.c_double(), C.c_double()
        native.check("jp_time_cmm", self.lib.jp_time_cmm(self.ctx, int(reps), C.byref(a), C.byref(b)))
        return a.value, b.value

    def get_lambda(self):
        """(lambda, its velocity)."""
        lam, lv = C.c_double(), C.c_double()
        native.check("jp_get_lambda", self.lib.jp_get_lambda(self.ctx, C.byref(lam), C.byref(lv)))
        return lam.value, lv.value

    def set_lambda(self, lam: float, lam_v: float = 0.0):
        native.check("jp_set_lambda", self.lib.jp_set_lambda(self.ctx, float(lam), float(lam_v)))

    def set_canonical(self, temps: Optional[float] = 300.0, icslop: int = 3, fcstol: float = 1e-3,
                      tempco: bool = False):
        """METHOD= CANO for the following runmd calls (SETTEM = temps; TEMPCO=
        YES with tempco: one temperature per molecule kind and cluster);
        temps=None returns to METHOD= MICR."""
        self._grp = None
        if tempco and temps is not None:
            self._grp = np.ascontiguousarray(self.system.temperature_groups() + 1, np.int32)
        on = 0 if temps is None else 1
        native.check("jp_set_canonical", self.lib.jp_set_canonical(
            self.ctx, on, float(temps or 0.0), int(icslop), float(fcstol),
            None if self._grp is None else _i32(self._grp)))

    def get_option(self, option: int) -> int:
        v = C.c_int32()
        native.check("jp_get_option", self.lib.jp_get_option(self.ctx, option, C.byref(v)))
        return v.value

    def timing(self):
        t = [C.c_double() for _ in range(4)]
        self.lib.jp_last_timing(self.ctx, *[C.byref(x) for x in t])
        return dict(total=t[0].value, pair=t[1].value, list=t[2].value, step=t[3].value)

    def time_pair(self, reps: int = 200) -> float:
        """Mean force-step pair-kernel duration (s) over reps back-to-back launches."""
        t = C.c_double()
        native.check("jp_time_pair", self.lib.jp_time_pair(self.ctx, self.precision, reps, C.byref(t)))
        return t.value

    def count_pairs(self, detail: bool = False):
        """Unique pairs with r <= rc; with detail, (all, both-LJ, listed)."""
        n, nlj, nl = C.c_int64(), C.c_int64(), C.c_int64()
        native.check("jp_count_pairs", self.lib.jp_count_pairs(self.ctx, C.byref(n), C.byref(nlj), C.byref(nl)))
        return (n.value, nlj.value, nl.value) if detail else n.value


def equilibrate(md: "MdSystem", nsteps: int, temp: float = 300.0, every: int = 10, update: int = 10):
    """Set-up helper (not part of the reference path): leap-frog steps with
    the velocities rescaled to `temp` every `every` steps, to start NVE runs
    from an equilibrated state."""
    from .systems import GASCST
    m = md.h.mass[:, None]
    ndf = md.system.ndf() - 3
    for _ in range(max(1, nsteps // every)):
        _, ier = md.runmd(every, update=update)
        if ier:
            return ier
        x, v = md.get_state()
        t = (m * v * v).sum() * 1e7 / (GASCST * ndf)
        md.set_state(None, v * np.sqrt(temp / t))
    return 0


def equilibrate_ensemble(ens: "Ensemble", nsteps: int, temp: float = 300.0, every: int = 10, update: int = 10):
    """equilibrate() for every system of a batch: batched leap-frog steps,
    each system's velocities rescaled to `temp` every `every` steps."""
    from .systems import GASCST
    for _ in range(max(1, nsteps // every)):
        _, ier = ens.runmd(every, update=update)
        if ier:
            return ier
        for i, h in enumerate(ens.holders):
            _, v = ens.get_state(i)
            t = (h.mass[:, None] * v * v).sum() * 1e7 / (GASCST * (h.system.ndf() - 3))
            ens.set_state(i, None, v * np.sqrt(temp / t))
    return 0


def _alsd_call(fn, name, handle, al):
    if al is None:
        native.check(name, fn(handle, 0, 0, 0, None, None, None, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0))
        return None
    c = np.ascontiguousarray(al["c"], np.float64)
    low = np.ascontiguousarray(al["low"], np.float64)
    high = np.ascontiguousarray(al["high"], np.float64)
    native.check(name, fn(handle, 1, c.shape[0], c.shape[1] - 1, _f64(c), _f64(low), _f64(high),
                          float(al["lower"]), float(al["upper"]), float(al["celw"]), float(al["ceup"]),
                          float(al["tempce"]), float(al.get("lmass", 0.0)), int(bool(al.get("tctl", True)))))
    return (c, low, high)


def log_to_array(log, nl: int) -> np.ndarray:
    dt = np.dtype([("step", "i4"), ("etot", "f8"), ("epot", "f8"), ("ekin", "f8"),
                   ("eelec", "f8", 3), ("evdw", "f8", 3), ("ebond", "f8", 3), ("eangle", "f8", 3),
                   ("etors", "f8", 3), ("eimp", "f8", 3), ("evdw14", "f8", 3), ("eelec14", "f8", 3),
                   ("lambda", "f8"), ("lambda_v", "f8")])
    out = np.zeros(nl, dt)
    for i in range(nl):
        L = log[i]
        out[i] = (L.step, L.etot, L.epot, L.ekin, tuple(L.eelec), tuple(L.evdw), tuple(L.ebond), tuple(L.eangle),
                  tuple(L.etors), tuple(L.eimp), tuple(L.evdw14), tuple(L.eelec14), L.lambda_, L.lambda_v)
    return out


class Ensemble:
    """Many independent systems of uneven size advanced by one set of
    batched launches (replaces one md_run process per system)."""

    def __init__(self, systems: Sequence[System], dt: float = 2.0, precision: int = FP32, **desc_kw):
        self.lib = native.load()
        self.holders = [DescHolder(s, dt=dt, **desc_kw) for s in systems]
        arr = (SysDesc * len(systems))(*[h.desc for h in self.holders])
        self._arr = arr
        self.systems = list(systems)
        self.nsys = len(systems)
        self.precision = precision
        p = C.c_void_p()
        native.check("jp_batch_create", self.lib.jp_batch_create(arr, self.nsys, C.byref(p)))
        self.b = p
        for i, s in enumerate(systems):
            v = np.ascontiguousarray(s.v, np.float64)
            native.check("jp_batch_set_state", self.lib.jp_batch_set_state(self.b, i, None, _f64(v)))

    def close(self):
        if getattr(self, "b", None):
            self.lib.jp_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def runmd(self, nsteps: int, update: int = 10, nmon: int = 0, precision: Optional[int] = None):
        prec = self.precision if precision is None else precision
        nl = nsteps // nmon + 1 if nmon > 0 else 0
        log = (EnergyLog * max(nl * self.nsys, 1))()
        ier = native.check("jp_batch_run_nve", self.lib.jp_batch_run_nve(self.b, nsteps, update, nmon, prec, log))
        full = log_to_array(log, nl * self.nsys)
        if not nl:
            return full, ier
        full = full.reshape(self.nsys, nl)
        used = full[0]["step"] >= 0        # every system records the same steps
        return full[:, used], ier

    def set_step(self, idloop: int):
        native.check("jp_batch_set_step", self.lib.jp_batch_set_step(self.b, int(idloop)))

    def restart_step(self, precision: Optional[int] = None):
        """MdSystem.restart_step for every system of the batch."""
        prec = self.precision if precision is None else precision
        log = (EnergyLog * self.nsys)()
        ier = native.check("jp_batch_restart_step", self.lib.jp_batch_restart_step(self.b, prec, log))
        return log_to_array(log, self.nsys), ier

    def prepare_graphs(self, nsteps: int, update: int = 10, nmon: int = 0, precision: Optional[int] = None):
        prec = self.precision if precision is None else precision
        native.check("jp_batch_prepare_graphs", self.lib.jp_batch_prepare_graphs(self.b, nsteps, update, nmon, prec))

    def set_graphs(self, on: bool):
        native.check("jp_batch_set_option", self.lib.jp_batch_set_option(self.b, native.JP_OPT_GRAPHS, int(bool(on))))

    def set_option(self, option: int, value: int):
        native.check("jp_batch_set_option", self.lib.jp_batch_set_option(self.b, option, int(value)))

    def set_alsd(self, al: Optional[dict]):
        """ALSD for every system of the batch (see MdSystem.set_alsd)."""
        self._al = _alsd_call(self.lib.jp_batch_set_alsd, "jp_batch_set_alsd", self.b, al)

    def get_lambda(self, i: int):
        lam, lv = C.c_double(), C.c_double()
        native.check("jp_batch_get_lambda", self.lib.jp_batch_get_lambda(self.b, i, C.byref(lam), C.byref(lv)))
        return lam.value, lv.value

    def set_lambda(self, i: int, lam: float, lam_v: float = 0.0):
        native.check("jp_batch_set_lambda", self.lib.jp_batch_set_lambda(self.b, i, float(lam), float(lam_v)))

    def set_stopce(self, mode: str, nstpcn: int = 1):
        """MdSystem.set_stopce for every system of the batch."""
        on = mode != "NO"
        self._stop = (np.array([s.chain_end(nstpcn) if on else 0 for s in self.systems], np.int32),
                      np.array([s.first_kind_atoms() for s in self.systems], np.int32))
        native.check("jp_batch_set_stopce", self.lib.jp_batch_set_stopce(
            self.b, STOPCE[mode], _i32(self._stop[0]), _i32(self._stop[1])))

    def start_step(self, precision: Optional[int] = None) -> int:
        """MdSystem.start_step for every system of the batch."""
        prec = self.precision if precision is None else precision
        return native.check("jp_batch_start_step", self.lib.jp_batch_start_step(self.b, prec))

    def set_canonical(self, temps: Optional[float] = 300.0, icslop: int = 3, fcstol: float = 1e-3,
                      tempco: bool = False):
        """METHOD= CANO for every system of the batch (see MdSystem.set_canonical)."""
        self._grp = None
        if tempco and temps is not None:
            self._grp = np.ascontiguousarray(
                np.concatenate([s.temperature_groups() + 1 for s in self.systems]), np.int32)
        on = 0 if temps is None else 1
        native.check("jp_batch_set_canonical", self.lib.jp_batch_set_canonical(
            self.b, on, float(temps or 0.0), int(icslop), float(fcstol),
            None if self._grp is None else _i32(self._grp)))

    def get_option(self, option: int) -> int:
        v = C.c_int32()
        native.check("jp_batch_get_option", self.lib.jp_batch_get_option(self.b, option, C.byref(v)))
        return v.value

    def get_state(self, i: int):
        n = self.holders[i].n
        x, v = np.zeros((n, 3)), np.zeros((n, 3))
        native.check("jp_batch_get_state", self.lib.jp_batch_get_state(self.b, i, _f64(x), _f64(v)))
        return x, v

    def set_state(self, i: int, cord=None, vel=None):
        cord = None if cord is None else np.ascontiguousarray(cord, np.float64)
        vel = None if vel is None else np.ascontiguousarray(vel, np.float64)
        native.check("jp_batch_set_state", self.lib.jp_batch_set_state(self.b, i, _f64(cord), _f64(vel)))

    def timing(self) -> float:
        t = C.c_double()
        self.lib.jp_batch_last_timing(self.b, C.byref(t))
        return t.value
