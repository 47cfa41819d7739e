"""ctypes mirror of include/jpgpu.h and the System -> jp_sys_desc builder.

The descriptor carries exactly what md_run's INPUT phase and incel derive from
the .tpl/.shk/md.inp files (see the field comments in include/jpgpu.h).
"""
from __future__ import annotations

import ctypes as C
from typing import List, Sequence

import numpy as np

from .systems import System

i32p = C.POINTER(C.c_int32)
f64p = C.POINTER(C.c_double)


class SysDesc(C.Structure):
    _fields_ = [
        ("natom", C.c_int32), ("ntyp", C.c_int32), ("nfrag", C.c_int32), ("nfrg", C.c_int32),
        ("x", f64p), ("mass", f64p), ("charge", f64p), ("atype", i32p),
        ("rmin2", f64p), ("eps", f64p), ("icls", i32p), ("n_matrix", i32p),
        ("nexcl", C.c_int32), ("excl", i32p), ("nexcess", C.c_int32), ("excess", i32p),
        ("nshake", C.c_int32), ("shake_natom", i32p), ("shake_atoms", i32p), ("shake_dist2", f64p),
        ("shake_tol", C.c_double), ("shake_maxit", C.c_int32), ("cell_level", C.c_int32),
        ("celwal", C.c_double * 6), ("cutoff", C.c_double), ("diel", C.c_double), ("dt", C.c_double),
        ("lmd", C.c_int32), ("lambda_", C.c_double),
    ]


class EnergyLog(C.Structure):
    _fields_ = [("step", C.c_int32), ("etot", C.c_double), ("epot", C.c_double), ("ekin", C.c_double),
                ("eelec", C.c_double * 3), ("evdw", C.c_double * 3)]


def n_matrix(nfrag: int, method: str) -> np.ndarray:
    """Cluster-pair -> energy-class table (reference inpcls.f90:297-324),
    returned column-major (Fortran order) as int32."""
    m = np.ones((nfrag, nfrag), np.int32)
    if nfrag == 2:
        if method.startswith("LMD"):
            m[:] = [[1, 2], [2, 3]]
        elif method == "AND":
            m[:] = 2
            m[0, 0] = 1
        elif method == "OR":
            m[:] = 1
            m[1, 1] = 2
        elif method == "XOR":
            m[:] = [[2, 1], [1, 2]]
    elif nfrag > 2:
        raise ValueError("nfrag > 2 not supported")
    return np.asfortranarray(m)


class DescHolder:
    """A SysDesc plus the numpy arrays it points into (kept alive)."""

    def __init__(self, s: System, dt: float = 2.0, cell_level: int = 3, shake_tol: float = 1e-6,
                 shake_maxit: int = 1000, x: np.ndarray = None):
        n = s.n
        self.system = s
        self.x = np.ascontiguousarray(s.x if x is None else x, np.float64)
        self.mass = np.ascontiguousarray(s.mass)
        self.charge = np.ascontiguousarray(s.charge)
        self.atype = np.ascontiguousarray(s.atype, np.int32)
        self.rmin2 = np.array([t.rmin2 for t in s.types], np.float64)
        self.eps = np.array([t.eps for t in s.types], np.float64)
        self.icls = np.ascontiguousarray(s.cluster, np.int32)
        nfrag = int(self.icls.max())
        self.nmat = np.ascontiguousarray(n_matrix(nfrag, s.cluster_method).ravel(order="F"), np.int32)
        nfrg = int(self.nmat.max())
        self.excl = np.ascontiguousarray(s.excl14() + 1, np.int32)
        self.excess = np.ascontiguousarray(s.excess_pairs() + 1, np.int32)
        groups = s.shake_groups()
        ng = len(groups)
        self.shk_n = np.array([len(a) for a, _ in groups], np.int32)
        self.shk_a = np.zeros((max(ng, 1), 4), np.int32)
        self.shk_d = np.zeros((max(ng, 1), 6), np.float64)
        for g, (atoms, d2) in enumerate(groups):
            self.shk_a[g, :len(atoms)] = np.array(atoms) + 1
            self.shk_d[g, :len(d2)] = d2
        d = SysDesc()
        d.natom, d.ntyp, d.nfrag, d.nfrg = n, len(s.types), nfrag, nfrg
        d.x = self.x.ctypes.data_as(f64p)
        d.mass = self.mass.ctypes.data_as(f64p)
        d.charge = self.charge.ctypes.data_as(f64p)
        d.atype = self.atype.ctypes.data_as(i32p)
        d.rmin2 = self.rmin2.ctypes.data_as(f64p)
        d.eps = self.eps.ctypes.data_as(f64p)
        d.icls = self.icls.ctypes.data_as(i32p)
        d.n_matrix = self.nmat.ctypes.data_as(i32p)
        d.nexcl = len(self.excl)
        d.excl = self.excl.ctypes.data_as(i32p)
        d.nexcess = len(self.excess)
        d.excess = self.excess.ctypes.data_as(i32p)
        d.nshake = ng
        d.shake_natom = self.shk_n.ctypes.data_as(i32p)
        d.shake_atoms = self.shk_a.ctypes.data_as(i32p)
        d.shake_dist2 = self.shk_d.ctypes.data_as(f64p)
        d.shake_tol = shake_tol
        d.shake_maxit = shake_maxit
        d.cell_level = cell_level
        for k, w in enumerate(s.celwal):
            d.celwal[k] = float(w)
        d.cutoff = s.cutoff
        d.diel = s.diel
        d.dt = dt
        d.lmd = 1 if s.cluster_method.startswith("LMD") else 0
        d.lambda_ = s.lam
        self.desc = d
        self.nfrg = nfrg
        self.n = n

    @property
    def ref(self):
        return C.byref(self.desc)
