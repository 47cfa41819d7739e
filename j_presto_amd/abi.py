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
        ("nbond", C.c_int32), ("bond_atoms", i32p), ("bond_fc", f64p), ("bond_r0", f64p), ("bond_cls", i32p),
        ("nangle", C.c_int32), ("angle_atoms", i32p), ("angle_fc", f64p), ("angle_theta0", f64p),
        ("angle_cls", i32p),
        ("ntors", C.c_int32), ("tors_atoms", i32p), ("tors_fc", f64p), ("tors_rot", f64p), ("tors_phase", f64p),
        ("tors_div", i32p), ("tors_cls", i32p),
        ("nimp", C.c_int32), ("imp_atoms", i32p), ("imp_fc", f64p), ("imp_rot", f64p), ("imp_phase", f64p),
        ("imp_div", i32p), ("imp_cls", i32p),
        ("n14", C.c_int32), ("p14_atoms", i32p), ("p14_vdw", f64p), ("p14_elec", f64p), ("p14_cls", i32p),
    ]


class EnergyLog(C.Structure):
    _fields_ = [("step", C.c_int32), ("etot", C.c_double), ("epot", C.c_double), ("ekin", C.c_double),
                ("eelec", C.c_double * 3), ("evdw", C.c_double * 3), ("ebond", C.c_double * 3),
                ("eangle", C.c_double * 3), ("etors", C.c_double * 3), ("eimp", C.c_double * 3),
                ("evdw14", C.c_double * 3), ("eelec14", C.c_double * 3), ("lambda_", C.c_double),
                ("lambda_v", C.c_double)]


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


# pairs of a SHAKE group's constraints, in the reference's order (setshk.f90 parlis)
_PARLIS = [(0, 1), (1, 2), (2, 0), (3, 0), (2, 3), (1, 3)]


def shake_removed_terms(bonds: np.ndarray, angles: np.ndarray, groups) -> tuple:
    """setshk.f90 reabnd / reaang: flags of the bonds and angles md_run drops
    because a SHAKE constraint holds their (1-2 or 1-3) distance.  The search
    follows the reference's loops: per group size, a forward scan from the
    position after the last match, abandoned once the term's atoms pass the
    constraint's.  `bonds` (nb,2) and `angles` (na,3) are 0-based in TPL
    order; `groups` is shake_groups() (sizes ascending).  Returns boolean
    arrays, True = removed."""
    off_b = np.zeros(len(bonds), bool)
    off_a = np.zeros(len(angles), bool)

    def passed(i1, i2, j1, j2):
        return (i1 < j1 and i2 < j2) or (i1 < j2 and i2 < j1)

    for size in (2, 3, 4):
        grp = [a for a, _ in groups if len(a) == size]
        nst = 0
        for atoms in grp:
            for p, q in _PARLIS[:size * (size - 1) // 2]:
                i1, i2 = atoms[p], atoms[q]
                for ib in range(nst, len(bonds)):
                    j1, j2 = int(bonds[ib, 0]), int(bonds[ib, 1])
                    if (i1 == j1 and i2 == j2) or (i1 == j2 and i2 == j1):
                        off_b[ib] = True
                        # 1-based: nst = ibnd + 1 (2-, 4-atom), max(ibnd - 1, 1) (3-atom)
                        nst = ib + 1 if size != 3 else max(ib, 1) - 1
                        break
                    if passed(i1, i2, j1, j2):
                        break
    for size in (3, 4):
        grp = [a for a, _ in groups if len(a) == size]
        nst = 0
        for atoms in grp:
            for p, q in _PARLIS[:size * (size - 1) // 2]:
                i1, i2 = atoms[p], atoms[q]
                for ia in range(nst, len(angles)):
                    j1, j2 = int(angles[ia, 0]), int(angles[ia, 2])
                    if (i1 == j1 and i2 == j2) or (i1 == j2 and i2 == j1):
                        off_a[ia] = True
                        # 1-based: nst = max(iang - 1, 1) (3-atom), max(iang - 2, 1) (4-atom)
                        nst = max(ia, 1) - 1 if size == 3 else max(ia - 1, 1) - 1
                        break
                    if passed(i1, i2, j1, j2):
                        break
    return off_b, off_a


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
        # bonded terms md_run keeps after setshk (reabnd / reaang), TPL order
        groups_all = s.shake_groups()
        bonds, angles = s.bonds(), s.angles()
        bfc, br0 = s.bond_params()
        afc, ath = s.angle_params()
        off_b, off_a = shake_removed_terms(bonds, angles, groups_all)
        keep_b = ~off_b
        keep_a = ~off_a
        self.bond_atoms = np.ascontiguousarray(bonds[keep_b] + 1, np.int32).reshape(-1, 2)
        self.bond_fc = np.ascontiguousarray(bfc[keep_b])
        self.bond_r0 = np.ascontiguousarray(br0[keep_b])
        self.angle_atoms = np.ascontiguousarray(angles[keep_a] + 1, np.int32).reshape(-1, 3)
        self.angle_fc = np.ascontiguousarray(afc[keep_a])
        self.angle_theta0 = np.ascontiguousarray(ath[keep_a] * (np.pi / 180.0))   # inptpl.f90 rdangl: * rad
        # a term with fc == 0 contributes nothing: leave such systems without bonded work
        if not np.any(self.bond_fc):
            self.bond_atoms, self.bond_fc, self.bond_r0 = self.bond_atoms[:0], self.bond_fc[:0], self.bond_r0[:0]
        if not np.any(self.angle_fc):
            self.angle_atoms, self.angle_fc = self.angle_atoms[:0], self.angle_fc[:0]
            self.angle_theta0 = self.angle_theta0[:0]
        d.nbond = len(self.bond_fc)
        d.bond_atoms = self.bond_atoms.ctypes.data_as(i32p)
        d.bond_fc = self.bond_fc.ctypes.data_as(f64p)
        d.bond_r0 = self.bond_r0.ctypes.data_as(f64p)
        d.nangle = len(self.angle_fc)
        d.angle_atoms = self.angle_atoms.ctypes.data_as(i32p)
        d.angle_fc = self.angle_fc.ctypes.data_as(f64p)
        d.angle_theta0 = self.angle_theta0.ctypes.data_as(f64p)
        # torsions incel keeps (|fc| >= epsilon, incel.f90:993-1009) and the 1-4
        # pairs of every torsion with the 1-4 flag (incel.f90:975-984); AMBER 1-4
        # scales min over the two types (inptpl.f90 mknopm), times ccoef q_i q_j
        tors = s.torsions()
        tfc, tdiv, trot, tph, tnbf = s.torsion_params()
        keep_t = np.abs(tfc) >= np.finfo(np.float64).eps
        self.tors_atoms = np.ascontiguousarray(tors[keep_t] + 1, np.int32).reshape(-1, 4)
        self.tors_fc = np.ascontiguousarray(tfc[keep_t], np.float64)
        self.tors_rot = np.ascontiguousarray(trot[keep_t], np.float64)
        self.tors_phase = np.ascontiguousarray(tph[keep_t] * (np.pi / 180.0), np.float64)
        self.tors_div = np.ascontiguousarray(tdiv[keep_t], np.int32)
        sel14 = tnbf == 1
        self.p14_atoms = np.ascontiguousarray(tors[sel14][:, [0, 3]] + 1, np.int32).reshape(-1, 2)
        sv = np.array([t_.scale14_vdw for t_ in s.types]); se = np.array([t_.scale14_elec for t_ in s.types])
        at = s.atype - 1
        a1, a4 = tors[sel14][:, 0], tors[sel14][:, 3]
        self.p14_vdw = np.ascontiguousarray(np.minimum(sv[at[a1]], sv[at[a4]]) * tnbf[sel14], np.float64)
        ess = np.minimum(se[at[a1]], se[at[a4]]) * tnbf[sel14]
        ccoef = 332.06378 / s.diel
        q = s.charge
        self.p14_elec = np.ascontiguousarray(ess * ccoef * q[a1] * q[a4], np.float64)
        d.ntors = len(self.tors_fc)
        d.tors_atoms = self.tors_atoms.ctypes.data_as(i32p)
        d.tors_fc = self.tors_fc.ctypes.data_as(f64p)
        d.tors_rot = self.tors_rot.ctypes.data_as(f64p)
        d.tors_phase = self.tors_phase.ctypes.data_as(f64p)
        d.tors_div = self.tors_div.ctypes.data_as(i32p)
        d.n14 = len(self.p14_vdw)
        d.p14_atoms = self.p14_atoms.ctypes.data_as(i32p)
        d.p14_vdw = self.p14_vdw.ctypes.data_as(f64p)
        d.p14_elec = self.p14_elec.ctypes.data_as(f64p)
        self.desc = d
        self.nfrg = nfrg
        self.n = n

    @property
    def ref(self):
        return C.byref(self.desc)
