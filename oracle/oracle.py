"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker; never by the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from j_presto_amd.abi import DescHolder

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None

i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
f64p = C.POINTER(C.c_double)


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.jo_create.restype = C.c_void_p
        L.jo_create.argtypes = [C.c_void_p]
        L.jo_destroy.argtypes = [C.c_void_p]
        L.jo_celprp.argtypes = [C.c_void_p, f64p]
        L.jo_get_list.argtypes = [C.c_void_p, i32p, i64p, i32p, i32p, C.c_int64]
        L.jo_get_order.argtypes = [C.c_void_p, i32p]
        L.jo_nonbond.argtypes = [C.c_void_p, f64p, f64p, f64p, f64p]
        L.jo_shake.argtypes = [C.c_void_p, f64p, f64p]
        L.jo_bonded.argtypes = [C.c_void_p, f64p, f64p, f64p]
        L.jo_verlet_step.argtypes = [C.c_void_p, f64p, f64p, f64p, f64p]
        L.jo_combine_grad.argtypes = [C.c_void_p, f64p, f64p]
        L.jo_cano_step.argtypes = [C.c_void_p, f64p, f64p, f64p, f64p, f64p, C.c_double, C.c_int, C.c_double,
                                   C.c_int, i32p, f64p]
        L.jo_group_dof.argtypes = [C.c_void_p, C.c_int, i32p, f64p]
        L.jo_ekin.argtypes = [C.c_void_p, f64p]
        L.jo_ekin.restype = C.c_double
        L.jo_run_nve.argtypes = [C.c_void_p, f64p, f64p, C.c_int, C.c_int]
        L.jo_run_cano.argtypes = [C.c_void_p, f64p, f64p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int,
                                  C.c_double, C.c_int, i32p, f64p, C.POINTER(GeParams), i32p]
        L.jo_run_cano_heat.argtypes = [C.c_void_p, f64p, f64p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int,
                                       C.c_double, C.c_int, i32p, f64p, C.c_int, C.c_int, C.c_double, i32p]
        L.jo_ge_weight.argtypes = [C.POINTER(GeParams), C.c_double]
        L.jo_ge_weight.restype = C.c_double
        L.jo_alsd_dew.argtypes = [C.POINTER(AlsdParams), C.c_double]
        L.jo_alsd_dew.restype = C.c_double
        L.jo_alsd_force.argtypes = [C.POINTER(AlsdParams), C.c_double, C.c_double, C.c_double]
        L.jo_alsd_force.restype = C.c_double
        L.jo_run_alsd.argtypes = [C.c_void_p, f64p, f64p, f64p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int,
                                  C.c_double, C.c_int, i32p, f64p, C.POINTER(AlsdParams), f64p, i32p]
        L.jo_cmm_setup.argtypes = [C.c_void_p, f64p, C.c_double, C.c_int]
        L.jo_cmm_celprp.argtypes = [C.c_void_p, f64p]
        L.jo_cmm_nonbond.argtypes = [C.c_void_p, f64p, f64p, f64p, f64p, C.c_int]
        L.jo_cmm_run_nve.argtypes = [C.c_void_p, f64p, f64p, C.c_int, C.c_int, C.c_int]
        L.jo_set_stopce.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.jo_start_step.argtypes = [C.c_void_p, f64p, f64p]
        L.jo_remotv.argtypes = [C.c_void_p, f64p, f64p]
        L.jo_stop_dof.argtypes = [C.c_void_p, C.c_int, i32p, C.c_int, f64p]
        L.jo_lambda.argtypes = [C.c_void_p]
        L.jo_lambda.restype = C.c_double
        L.jo_set_lambda.argtypes = [C.c_void_p, C.c_double]
        L.jo_list_brute.argtypes = [C.c_void_p, f64p, i32p, i32p, C.c_int64]
        L.jo_list_brute.restype = C.c_int64
        L.jo_constants.argtypes = [C.c_void_p, f64p]
        L.jo_dself.argtypes = [C.c_void_p, f64p]
        L.jo_chgmod.argtypes = [C.c_void_p, f64p]
        L.jo_ljtable.argtypes = [C.c_void_p, f64p, f64p]
        _lib = L
    return _lib


def _p(a, t=f64p):
    return a.ctypes.data_as(t)


class GeParams(C.Structure):
    """jo_ge: multicanonical parameters (inmcmd.f90:20-31)."""
    _fields_ = [("nwindow", C.c_int), ("ndeg", C.c_int), ("c", f64p), ("low", f64p), ("high", f64p),
                ("alphalw", C.c_double), ("alphaup", C.c_double), ("celw", C.c_double), ("ceup", C.c_double),
                ("tempce", C.c_double)]


def ge_params(ge: dict):
    """(GeParams, arrays it points into) from a dict with c (nwindow, ndeg+1),
    low, high (nwindow), alphalw, alphaup, celw, ceup, tempce."""
    c = np.ascontiguousarray(ge["c"], np.float64)
    low = np.ascontiguousarray(ge["low"], np.float64)
    high = np.ascontiguousarray(ge["high"], np.float64)
    p = GeParams(c.shape[0], c.shape[1] - 1, _p(c), _p(low), _p(high), float(ge["alphalw"]), float(ge["alphaup"]),
                 float(ge["celw"]), float(ge["ceup"]), float(ge["tempce"]))
    return p, (c, low, high)


class AlsdParams(C.Structure):
    """jo_alsd: adaptive lambda square dynamics (inmcmd.f90:20-41 under
    GEPSME= LMD, LWEIGH, LTCNTL)."""
    _fields_ = [("nwindow", C.c_int), ("ndeg", C.c_int), ("c", f64p), ("low", f64p), ("high", f64p),
                ("lower", C.c_double), ("upper", C.c_double), ("celw", C.c_double), ("ceup", C.c_double),
                ("tempce", C.c_double), ("lmass", C.c_double), ("tctl", C.c_int), ("lgrp", C.c_int)]


def alsd_params(al: dict, lgrp: int = -1):
    """(AlsdParams, arrays it points into) from a dict with c (nwindow,
    ndeg+1), low, high, lower, upper, celw, ceup, tempce, lmass, tctl."""
    c = np.ascontiguousarray(al["c"], np.float64)
    low = np.ascontiguousarray(al["low"], np.float64)
    high = np.ascontiguousarray(al["high"], np.float64)
    p = AlsdParams(c.shape[0], c.shape[1] - 1, _p(c), _p(low), _p(high), float(al["lower"]), float(al["upper"]),
                   float(al["celw"]), float(al["ceup"]), float(al["tempce"]), float(al["lmass"]),
                   int(bool(al["tctl"])), int(lgrp))
    return p, (c, low, high)


def alsd_dew(al: dict, lam: float) -> float:
    """ALSD's bias dew at lambda (the polynomial windows plus the walls)."""
    p, keep = alsd_params(al)
    return lib().jo_alsd_dew(C.byref(p), float(lam))


def ge_weight(ge: dict, e: float) -> float:
    """md_run's multicanonical force weight at class-1 energy e."""
    p, keep = ge_params(ge)
    return lib().jo_ge_weight(C.byref(p), float(e))


class Oracle:
    """FP64 CPU restatement of md_run's ZD path for one system."""

    def __init__(self, system, dt: float = 2.0, **kw):
        self.h = DescHolder(system, dt=dt, **kw)
        self.n = self.h.n
        self.nfrg = self.h.nfrg
        self.ntyp = len(system.types)
        self.c = lib().jo_create(C.byref(self.h.desc))
        self.stop_kind = 0

    def __del__(self):
        try:
            if self.c:
                lib().jo_destroy(self.c)
                self.c = None
        except Exception:
            pass

    def celprp(self, x) -> int:
        x = np.ascontiguousarray(x, np.float64)
        return lib().jo_celprp(self.c, _p(x))

    def get_list(self):
        npairs = C.c_int64()
        lib().jo_get_list(self.c, None, C.byref(npairs), None, None, 0)
        cell = np.zeros((self.n, 3), np.int32)
        ptr = np.zeros(self.n + 1, np.int32)
        idx = np.zeros(max(npairs.value, 1), np.int32)
        r = lib().jo_get_list(self.c, _p(cell, i32p), C.byref(npairs), _p(ptr, i32p), _p(idx, i32p), len(idx))
        assert r == 0
        return cell, ptr, idx[:npairs.value]

    def order(self):
        o = np.zeros(self.n, np.int32)
        lib().jo_get_order(self.c, _p(o, i32p))
        return o

    def nonbond(self, x, grad=None):
        x = np.ascontiguousarray(x, np.float64)
        g = np.zeros((self.nfrg, self.n, 3)) if grad is None else grad
        ee = np.zeros(self.nfrg)
        ev = np.zeros(self.nfrg)
        r = lib().jo_nonbond(self.c, _p(x), _p(g), _p(ee), _p(ev))
        assert r == 0
        return g, ee, ev

    def bonded(self, x, grad=None):
        """dnergv_PB's bonded part: (grad(nfrg,N,3) accumulated, eterm(nfrg,6))
        with eterm columns bond, angle, torsion, improper, 1-4 vdW, 1-4 elec."""
        x = np.ascontiguousarray(x, np.float64)
        g = np.zeros((self.nfrg, self.n, 3)) if grad is None else grad
        et = np.zeros((self.nfrg, 6))
        r = lib().jo_bonded(self.c, _p(x), _p(g), _p(et))
        assert r == 0
        return g, et

    def shake(self, x_new, x_old) -> int:
        return lib().jo_shake(self.c, _p(x_new), _p(np.ascontiguousarray(x_old)))

    def combine(self, grad):
        gt = np.zeros((self.n, 3))
        lib().jo_combine_grad(self.c, _p(np.ascontiguousarray(grad)), _p(gt))
        return gt

    def verlet_step(self, x, v, gtot) -> int:
        xo = np.zeros_like(x)
        return lib().jo_verlet_step(self.c, _p(x), _p(v), _p(np.ascontiguousarray(gtot)), _p(xo))

    def group_dof(self, grp=None):
        """Degrees of freedom per temperature group (grp: 0-based per atom, or
        None for one group)."""
        ng = 1 if grp is None else int(grp.max()) + 1
        dof = np.zeros(ng)
        g = None if grp is None else np.ascontiguousarray(grp, np.int32)
        lib().jo_group_dof(self.c, ng, None if g is None else _p(g, i32p), _p(dof))
        lib().jo_stop_dof(self.c, ng, None if g is None else _p(g, i32p), self.stop_kind, _p(dof))
        return dof

    def cano_step(self, x, v, gtot, temps: float, icslop: int = 3, fcstol: float = 1e-3, grp=None):
        """One METHOD= CANO step in place; returns (ier, vback)."""
        xo = np.zeros_like(x)
        vb = np.zeros_like(v)
        dof = self.group_dof(grp)
        g = None if grp is None else np.ascontiguousarray(grp, np.int32)
        ier = lib().jo_cano_step(self.c, _p(x), _p(v), _p(np.ascontiguousarray(gtot)), _p(xo), _p(vb), temps,
                                 icslop, fcstol, len(dof), None if g is None else _p(g, i32p), _p(dof))
        return ier, vb

    def ekin(self, v) -> float:
        return lib().jo_ekin(self.c, _p(np.ascontiguousarray(v)))

    def run_nve(self, x, v, nsteps: int, update: int = 10) -> int:
        return lib().jo_run_nve(self.c, _p(x), _p(v), nsteps, update)

    def run_cano(self, x, v, nsteps: int, update: int = 10, temps: float = 300.0, icslop: int = 3,
                 fcstol: float = 1e-3, grp=None, inirst: bool = True, ge=None, heat=None):
        """md_run's METHOD= CANO loop with verlethe's SHAKE retry (METHOD= GE
        with the multicanonical parameters ge), in place, after the restart
        set-up step when inirst; heat = (restart loop, HEATLO, STARTT) runs
        md_run's heating loop.  Returns (ier, provisions)."""
        dof = self.group_dof(grp)
        g = None if grp is None else np.ascontiguousarray(grp, np.int32)
        nprov = C.c_int32()
        if heat is not None:
            assert ge is None
            loop0, htlp, tmpsta = heat
            ier = lib().jo_run_cano_heat(self.c, _p(x), _p(v), nsteps, update, int(inirst), temps, icslop, fcstol,
                                         len(dof), None if g is None else _p(g, i32p), _p(dof), int(loop0),
                                         int(htlp), float(tmpsta), C.byref(nprov))
            return ier, nprov.value
        gp, keep = ge_params(ge) if ge is not None else (None, None)
        ier = lib().jo_run_cano(self.c, _p(x), _p(v), nsteps, update, int(inirst), temps, icslop, fcstol, len(dof),
                                None if g is None else _p(g, i32p), _p(dof),
                                None if gp is None else C.byref(gp), C.byref(nprov))
        return ier, nprov.value

    @property
    def lam(self) -> float:
        return lib().jo_lambda(self.c)

    @lam.setter
    def lam(self, value: float):
        lib().jo_set_lambda(self.c, float(value))

    def alsd_dof(self, grp=None, tctl: bool = True):
        """Degrees of freedom per temperature group with lambda's (md.f90:
        114-123 for one group; 221-228 under TEMPCO= YES: the first group with
        the most degrees of freedom takes it), and that group (-1: one group
        or no lambda temperature control)."""
        dof = self.group_dof(grp)
        lgrp = -1
        if tctl:
            if grp is None:
                dof[0] += 1.0
            else:
                lgrp = int(np.argmax(dof))
                dof[lgrp] += 1.0
        return dof, lgrp

    def run_alsd(self, x, v, lam_v: float, nsteps: int, alsd: dict, update: int = 10, temps: float = 300.0,
                 icslop: int = 3, fcstol: float = 1e-3, grp=None, inirst: bool = True):
        """md_run's ALSD loop (METHOD= GE, GEPSME= LMD) in place from the
        oracle's lambda and lam_v; returns (ier, lambda, lambda_v, lambda after
        every step, provisions)."""
        dof, lgrp = self.alsd_dof(grp, bool(alsd["tctl"]))
        g = None if grp is None else np.ascontiguousarray(grp, np.int32)
        ap, keep = alsd_params(alsd, lgrp)
        lv = C.c_double(float(lam_v))
        lam_log = np.zeros(max(nsteps, 1))
        nprov = C.c_int32()
        ier = lib().jo_run_alsd(self.c, _p(x), _p(v), C.byref(lv), nsteps, update, int(inirst), temps, icslop,
                                fcstol, len(dof), None if g is None else _p(g, i32p), _p(dof), C.byref(ap),
                                _p(lam_log), C.byref(nprov))
        return ier, self.lam, lv.value, lam_log[:nsteps], nprov.value

    # ---- CAL15M= CMM (non-periodic cell multipole method) ----------------
    def cmm_setup(self, x, cminsiz: float = 6.0, nlev: int = 3) -> int:
        """incel's CMM set-up (CELSIZ, CELLVL; nlev raised to the extent of x)."""
        return lib().jo_cmm_setup(self.c, _p(np.ascontiguousarray(x, np.float64)), cminsiz, nlev)

    def cmm_celprp(self, x) -> int:
        return lib().jo_cmm_celprp(self.c, _p(np.ascontiguousarray(x, np.float64)))

    def cmm_nonbond(self, x, grad=None, itwin: bool = True):
        """cmmdo under CMM: near tables and far field; returns (grad, eelec,
        evdw) with grad accumulated (nfrg, N, 3)."""
        if grad is None:
            grad = np.zeros((self.nfrg, self.n, 3))
        ee = np.zeros(self.nfrg)
        ev = np.zeros(self.nfrg)
        ier = lib().jo_cmm_nonbond(self.c, _p(np.ascontiguousarray(x, np.float64)), _p(grad), _p(ee), _p(ev),
                                   int(bool(itwin)))
        assert ier == 0, ier
        return grad, ee, ev

    def cmm_run_nve(self, x, v, nsteps: int, update: int = 10, cmmupd: int = 5) -> int:
        return lib().jo_cmm_run_nve(self.c, _p(x), _p(v), nsteps, update, cmmupd)

    STOPCE = {"NO": 0, "TRAN": 1, "ROTA": 2, "BOTH": 3}

    def set_stopce(self, mode: str, natom: int, nkind: int = 0) -> int:
        """STOPCE= TRAN/ROTA/BOTH/NO over the first natom atoms (md_run's
        ixcend(NSTPCN)) for every later step; nkind: atoms from atom 1 of its
        molecule kind (the degrees of freedom under TEMPCO= YES)."""
        self.stop_kind = int(nkind)
        return lib().jo_set_stopce(self.c, self.STOPCE[mode], int(natom))

    def remotv(self, xp, v) -> int:
        """remotv alone: the stopped motions taken out of v in place."""
        return lib().jo_remotv(self.c, _p(np.ascontiguousarray(xp, np.float64)), _p(v))

    def start_step(self, x, v) -> int:
        """md_run's initial step of a run without a restart (staver), in
        place."""
        return lib().jo_start_step(self.c, _p(x), _p(v))

    def list_brute(self, x):
        x = np.ascontiguousarray(x, np.float64)
        n = lib().jo_list_brute(self.c, _p(x), None, None, 0)
        ptr = np.zeros(self.n + 1, np.int32)
        idx = np.zeros(max(n, 1), np.int32)
        lib().jo_list_brute(self.c, _p(x), _p(ptr, i32p), _p(idx, i32p), len(idx))
        return ptr, idx[:n]

    def constants(self):
        o = np.zeros(16)
        lib().jo_constants(self.c, _p(o))
        keys = ["zcore", "bcoeff", "fcoeff", "irlim2", "rlim", "rlim2", "siz", "org_x", "org_y", "org_z",
                "nv", "nlev", "mx1", "mx2", "mxdelay", "ccoef"]
        return dict(zip(keys, o))

    def dself(self):
        o = np.zeros(self.nfrg)
        lib().jo_dself(self.c, _p(o))
        return o

    def chgmod(self):
        o = np.zeros(self.n)
        lib().jo_chgmod(self.c, _p(o))
        return o

    def ljtable(self):
        c6 = np.zeros((self.ntyp, self.ntyp))
        c12 = np.zeros((self.ntyp, self.ntyp))
        lib().jo_ljtable(self.c, _p(c6), _p(c12))
        return c6, c12
