"""Synthetic periodic systems in j_presto's data model.

A system is what md_run reads from its .tpl (topology), .shk (SHAKE groups),
coordinate file and restart file: molecules with per-copy atom records, LJ
atom types (AMBER Rmin/2 + epsilon), bonds/angles that define the excluded
pairs, SHAKE groups with squared target distances, cluster classes, and FP64
coordinates/velocities.  The same object feeds three consumers: the TPL/SHK/
restart writers (so the compiled reference can run the exact system), the CPU
oracle, and the GPU library through its C ABI.

Geometry and parameters of the workloads:
  * rigid TIP3P-like water: O-H 0.9572 A, H-O-H 104.52 deg, q(O)=-0.834,
    q(H)=+0.417, LJ on O only (sigma 3.1507 A, eps 0.1521 kcal/mol),
    O on a jittered cubic lattice (U[-0.2,0.2] A), uniform random orientation;
  * Na+/Cl- with Joung-Cheatham TIP3P LJ;
  * a self-avoiding random-walk chain, 1.5 A bonds, charges N(0,0.3) shifted
    to net zero, LJ sigma U[1.9,3.8] A, eps U[0.015,0.2] kcal/mol drawn per
    chain atom type, every 4th atom an "H" SHAKE-constrained to its
    predecessor (independent 2-atom groups).
Seeds are numpy PCG64 with seed = 1000 + workload index.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional, Sequence, Tuple

import numpy as np

# Physical constants (reference PHYCNS.f90:20-31)
GASCST = 8.31451
JOUCAL = 4.184
EEKCAL = 332.06378

R_OH = 0.9572
THETA_HOH = 104.52
R_HH = 2.0 * R_OH * math.sin(math.radians(THETA_HOH) / 2.0)
SIG_O, EPS_O = 3.1507, 0.1521
RMIN2_O = 2.0 ** (1.0 / 6.0) * SIG_O / 2.0
MASS_O, MASS_H = 15.9994, 1.008
Q_O, Q_H = -0.834, 0.417
# Joung & Cheatham (2008) TIP3P-specific ion parameters (Rmin/2, eps)
NA_RMIN2, NA_EPS, NA_MASS = 1.212, 0.3526418, 22.98977
CL_RMIN2, CL_EPS, CL_MASS = 2.513, 0.0355910, 35.453
WATER_PER_A3 = 0.997 / 18.01528 * 6.02214076e23 / 1e24   # 0.997 g/cm^3
BULK_WATER_PER_A3 = 2000.0 / 39.1 ** 3                     # the 6k-atom water box: 1.0007 g/cm^3
# water volume a compact-chain atom displaces in the "bulk" solvent mode
# (measured on chains of 300-3,000 atoms: 26.3-29.2 A^3), used to size boxes
CHAIN_ATOM_A3 = 28.0


@dataclasses.dataclass
class MolTemplate:
    """One TPL molecule (TPL> MOLECULES entry) and its per-copy records."""
    name: str
    ncopy: int
    aname: List[str]
    atname: List[str]
    itype: np.ndarray            # (na,) 1-based LJ type
    rname: List[str]
    resnum: np.ndarray           # (na,) residue number within the copy
    mass: np.ndarray
    charge: np.ndarray
    n1x: np.ndarray              # (na,3) counts of forward 1-2/1-3/1-4 partners
    partners: List[List[int]]    # forward partner offsets (> 0)
    bonds: np.ndarray            # (nb,2) local 1-based
    bond_fc: np.ndarray
    bond_r0: np.ndarray
    angles: np.ndarray           # (na,3) local 1-based
    angle_fc: np.ndarray
    angle_deg: np.ndarray
    shake: List[Tuple[Tuple[int, ...], Tuple[float, ...]]]  # local 1-based atoms, distances
    cluster: int = 1             # cluster class of every atom of this molecule
    # TPL> TORSIONS records: atoms (nt,4) local 1-based, force constant, division
    # number, periodicity, phase (deg), 1-4 flag (iytnbf)
    torsions: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0, 4), int))
    tors_fc: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    tors_div: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, int))
    tors_rot: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    tors_phase_deg: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    tors_nbf: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, int))

    @property
    def natom(self) -> int:
        return len(self.aname)


@dataclasses.dataclass
class AtomType:
    name: str
    rmin2: float     # AMBER Rmin/2 (TPL NONB column 4, fyvwrd)
    eps: float       # well depth (fyvwme)
    scale14_vdw: float = 0.5                  # fy14sv (TPL NONB column 7)
    scale14_elec: float = 0.8333333333333334  # fy14se (TPL NONB column 6)


@dataclasses.dataclass
class System:
    name: str
    mols: List[MolTemplate]
    types: List[AtomType]
    box: np.ndarray              # (3,) edge lengths; the box is [0, L)
    x: np.ndarray                # (N,3) FP64
    v: np.ndarray                # (N,3) FP64, velocity at t - dt/2 (restart semantics)
    cutoff: float = 12.0
    diel: float = 1.0
    cluster_method: str = "AND"
    lam: float = 0.0

    # ---- derived, atom-expanded arrays -------------------------------
    def _expand(self, attr):
        out = []
        for m in self.mols:
            a = getattr(m, attr)
            out.append(np.tile(np.asarray(a), m.ncopy) if np.ndim(a) == 1 else np.tile(a, (m.ncopy, 1)))
        return np.concatenate(out)

    @property
    def n(self) -> int:
        return sum(m.natom * m.ncopy for m in self.mols)

    @property
    def mass(self) -> np.ndarray:
        return self._expand("mass").astype(np.float64)

    @property
    def charge(self) -> np.ndarray:
        return self._expand("charge").astype(np.float64)

    @property
    def atype(self) -> np.ndarray:
        return self._expand("itype").astype(np.int32)

    @property
    def cluster(self) -> np.ndarray:
        return np.concatenate([np.full(m.natom * m.ncopy, m.cluster, np.int32) for m in self.mols])

    @property
    def nfrag(self) -> int:
        return int(self.cluster.max())

    def temperature_groups(self) -> np.ndarray:
        """0-based temperature group of every atom for TEMPCO= YES (md.f90:
        136-152): ixamol2 = (molecule kind - 1) * nfrag + cluster, renumbered
        without the empty values."""
        kind = np.concatenate([np.full(m.natom * m.ncopy, k + 1, np.int64) for k, m in enumerate(self.mols)])
        g = (kind - 1) * self.nfrag + self.cluster
        return np.unique(g, return_inverse=True)[1].astype(np.int32)

    def chain_end(self, nstpcn: int = 1) -> int:
        """Atoms of the first nstpcn chains (molecule copies): md_run's
        ixcend(nstpcn) (inptpl.f90:279, 680), the atoms STOPCE acts on."""
        ends = np.cumsum([m.natom for m in self.mols for _ in range(m.ncopy)])
        return int(ends[nstpcn - 1])

    def first_kind_atoms(self) -> int:
        """Atoms from atom 1 of atom 1's molecule kind (md.f90:187-193)."""
        return int(self.mols[0].natom * self.mols[0].ncopy)

    @property
    def celwal(self) -> np.ndarray:
        return np.array([0.0, self.box[0], 0.0, self.box[1], 0.0, self.box[2]])

    def _offsets(self):
        off = 0
        for m in self.mols:
            for c in range(m.ncopy):
                yield m, off
                off += m.natom

    def bonds(self) -> np.ndarray:
        out = [m.bonds - 1 + off for m, off in self._offsets() if len(m.bonds)]
        return np.concatenate(out).astype(np.int64) if out else np.zeros((0, 2), np.int64)

    def angles(self) -> np.ndarray:
        out = [m.angles - 1 + off for m, off in self._offsets() if len(m.angles)]
        return np.concatenate(out).astype(np.int64) if out else np.zeros((0, 3), np.int64)

    def torsions(self) -> np.ndarray:
        out = [m.torsions - 1 + off for m, off in self._offsets() if len(m.torsions)]
        return np.concatenate(out).astype(np.int64) if out else np.zeros((0, 4), np.int64)

    def torsion_params(self):
        """(fc, div, rot, phase_deg, nbf) of every torsion, in torsions() order."""
        keys = ("tors_fc", "tors_div", "tors_rot", "tors_phase_deg", "tors_nbf")
        parts = [[np.asarray(getattr(m, k)) for m, _ in self._offsets() if len(m.torsions)] for k in keys]
        if not parts[0]:
            return np.zeros(0), np.zeros(0, int), np.zeros(0), np.zeros(0), np.zeros(0, int)
        return tuple(np.concatenate(p) for p in parts)

    def bond_params(self):
        """(fc, r0) of every bond, in bonds() order."""
        fc = [np.asarray(m.bond_fc, np.float64) for m, _ in self._offsets() if len(m.bonds)]
        r0 = [np.asarray(m.bond_r0, np.float64) for m, _ in self._offsets() if len(m.bonds)]
        return (np.concatenate(fc), np.concatenate(r0)) if fc else (np.zeros(0), np.zeros(0))

    def angle_params(self):
        """(fc, theta0 in degrees) of every angle, in angles() order."""
        fc = [np.asarray(m.angle_fc, np.float64) for m, _ in self._offsets() if len(m.angles)]
        th = [np.asarray(m.angle_deg, np.float64) for m, _ in self._offsets() if len(m.angles)]
        return (np.concatenate(fc), np.concatenate(th)) if fc else (np.zeros(0), np.zeros(0))

    def excl14(self) -> np.ndarray:
        """Unique excluded pairs (i<j, 0-based) from the ATOM partner lists
        (the reference's itab14, inpcls.f90:229-295)."""
        out = []
        for m, off in self._offsets():
            for a, plist in enumerate(m.partners):
                for p in plist:
                    out.append((off + a, off + a + p))
        if not out:
            return np.zeros((0, 2), np.int64)
        e = np.array(out, np.int64)
        e.sort(axis=1)
        return np.unique(e, axis=0)

    def excess_pairs(self) -> np.ndarray:
        """ZD excess pairs: every bond, every angle's 1-3 pair, then the 1-4
        pair of every torsion with the 1-4 flag, in TPL order (reference
        incel.f90:665-688)."""
        b = self.bonds()
        a = self.angles()
        parts = [b[:, :2]] if len(b) else []
        if len(a):
            parts.append(a[:, [0, 2]])
        tr = self.torsions()
        if len(tr):
            nbf = self.torsion_params()[4]
            parts.append(tr[nbf == 1][:, [0, 3]])
        return np.concatenate(parts).astype(np.int64) if parts else np.zeros((0, 2), np.int64)

    def shake_groups(self):
        """List of (atoms 0-based tuple, squared distances tuple), ordered
        2-atom groups first, then 3-atom, then 4-atom (reference
        setshk.f90 modshk)."""
        g = []
        for m, off in self._offsets():
            for atoms, dist in m.shake:
                g.append((tuple(off + a - 1 for a in atoms), tuple(d * d for d in dist)))
        g.sort(key=lambda t: len(t[0]))  # stable: keeps input order within a size
        return g

    def ndf(self) -> int:
        ncon = sum(len(a) * (len(a) - 1) // 2 for a, _ in self.shake_groups())
        return 3 * self.n - ncon


# ---------------------------------------------------------------------------
# templates
# ---------------------------------------------------------------------------

def water_template(ncopy: int, type_o: int = 1, type_h: int = 2, cluster: int = 1) -> MolTemplate:
    return MolTemplate(
        name="WAT", ncopy=ncopy, aname=["O", "H1", "H2"], atname=["OW", "HW", "HW"],
        itype=np.array([type_o, type_h, type_h]), rname=["WAT"] * 3, resnum=np.array([1, 1, 1]),
        mass=np.array([MASS_O, MASS_H, MASS_H]), charge=np.array([Q_O, Q_H, Q_H]),
        n1x=np.array([[2, 0, 0], [0, 1, 0], [0, 0, 0]]), partners=[[1, 2], [1], []],
        bonds=np.array([[1, 2], [1, 3]]), bond_fc=np.array([553.0, 553.0]), bond_r0=np.array([R_OH, R_OH]),
        angles=np.array([[2, 1, 3]]), angle_fc=np.array([100.0]), angle_deg=np.array([THETA_HOH]),
        shake=[((1, 2, 3), (R_OH, R_HH, R_OH))], cluster=cluster)


def ion_template(name: str, ncopy: int, itype: int, mass: float, charge: float, cluster: int = 1) -> MolTemplate:
    return MolTemplate(
        name=name, ncopy=ncopy, aname=[name], atname=[name], itype=np.array([itype]), rname=[name],
        resnum=np.array([1]), mass=np.array([mass]), charge=np.array([charge]),
        n1x=np.zeros((1, 3), int), partners=[[]], bonds=np.zeros((0, 2), int), bond_fc=np.zeros(0),
        bond_r0=np.zeros(0), angles=np.zeros((0, 3), int), angle_fc=np.zeros(0), angle_deg=np.zeros(0),
        shake=[], cluster=cluster)


def water_types() -> List[AtomType]:
    return [AtomType("OW", RMIN2_O, EPS_O), AtomType("HW", 0.0, 0.0)]


# ---------------------------------------------------------------------------
# geometry helpers
# ---------------------------------------------------------------------------

def _water_body() -> np.ndarray:
    """O at origin, H's in the xy plane, bisector along +x."""
    h = math.radians(THETA_HOH) / 2.0
    return np.array([[0.0, 0.0, 0.0],
                     [R_OH * math.cos(h), R_OH * math.sin(h), 0.0],
                     [R_OH * math.cos(h), -R_OH * math.sin(h), 0.0]])


def _random_rotations(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform random rotation matrices from uniform unit quaternions (Shoemake)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a = np.sqrt(1 - u1)
    b = np.sqrt(u1)
    qw = a * np.sin(2 * np.pi * u2)
    qx = a * np.cos(2 * np.pi * u2)
    qy = b * np.sin(2 * np.pi * u3)
    qz = b * np.cos(2 * np.pi * u3)
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (qy * qy + qz * qz)
    R[:, 0, 1] = 2 * (qx * qy - qz * qw)
    R[:, 0, 2] = 2 * (qx * qz + qy * qw)
    R[:, 1, 0] = 2 * (qx * qy + qz * qw)
    R[:, 1, 1] = 1 - 2 * (qx * qx + qz * qz)
    R[:, 1, 2] = 2 * (qy * qz - qx * qw)
    R[:, 2, 0] = 2 * (qx * qz - qy * qw)
    R[:, 2, 1] = 2 * (qy * qz + qx * qw)
    R[:, 2, 2] = 1 - 2 * (qx * qx + qy * qy)
    return R


def _lattice_sites(nsite: int, L: float) -> np.ndarray:
    m = int(math.ceil(nsite ** (1.0 / 3.0) - 1e-9))
    sp = L / m
    g = (np.arange(m) + 0.5) * sp
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)


def _place_waters(rng, oxy: np.ndarray, L: np.ndarray, fixed: Optional[np.ndarray] = None,
                  min_sep: float = 1.6, iters: int = 40) -> np.ndarray:
    """Random orientations for waters with O at `oxy`; orientations that put
    any atom closer than min_sep to another molecule's atom are redrawn."""
    from scipy.spatial import cKDTree
    nw = len(oxy)
    body = _water_body()
    rot = _random_rotations(rng, nw)
    for _ in range(iters):
        pos = oxy[:, None, :] + np.einsum("nij,kj->nki", rot, body)
        flat = np.mod(pos.reshape(-1, 3), L)
        pts = flat if fixed is None else np.concatenate([flat, np.mod(fixed, L)])
        tree = cKDTree(pts, boxsize=L)
        pairs = tree.query_pairs(min_sep, output_type="ndarray")
        if len(pairs) == 0:
            break
        mol = np.where(pairs < 3 * nw, pairs // 3, -1)
        inter = mol[:, 0] != mol[:, 1]
        bad = np.unique(mol[inter].ravel())
        bad = bad[bad >= 0]
        if len(bad) == 0:
            break
        rot[bad] = _random_rotations(rng, len(bad))
    pos = oxy[:, None, :] + np.einsum("nij,kj->nki", rot, body)
    return pos.reshape(-1, 3)


def thermal_velocities(rng, sys: System, temp: float = 300.0) -> np.ndarray:
    """Maxwell-Boltzmann velocities (A/fs) consistent with the SHAKE groups:
    rigid-body projection for 3-atom groups, bond-component removal for
    2-atom groups, centre-of-mass motion removed, rescaled to `temp` with the
    reference's temperature formula (caltmp.f90: T = sum m v^2 1e7/(R ndf))."""
    m = sys.mass
    n = len(m)
    kT = GASCST * temp * 1e-7                     # (g/mol) (A/fs)^2
    v = rng.normal(size=(n, 3)) * np.sqrt(kT / m)[:, None]
    x = sys.x
    L = sys.box
    for atoms_list in _groups_by_size(sys):
        atoms = np.array(atoms_list)
        if atoms.shape[1] == 3:
            r = x[atoms]
            r = r - r[:, :1]
            r -= L * np.round(r / L)
            mm = m[atoms]
            M = mm.sum(1)
            com = (mm[:, :, None] * r).sum(1) / M[:, None]
            rr = r - com[:, None]
            vv = v[atoms]
            vcom = (mm[:, :, None] * vv).sum(1) / M[:, None]
            Lm = (mm[:, :, None] * np.cross(rr, vv - vcom[:, None])).sum(1)
            r2 = (rr * rr).sum(2)
            I = (mm * r2).sum(1)[:, None, None] * np.eye(3) - np.einsum("nk,nki,nkj->nij", mm, rr, rr)
            w = np.linalg.solve(I, Lm[:, :, None])[:, :, 0]
            v[atoms] = vcom[:, None] + np.cross(w[:, None, :], rr)
        elif atoms.shape[1] == 2:
            i, j = atoms[:, 0], atoms[:, 1]
            d = x[i] - x[j]
            d -= L * np.round(d / L)
            u = d / np.linalg.norm(d, axis=1)[:, None]
            vr = ((v[i] - v[j]) * u).sum(1)[:, None] * u
            mi, mj = m[i][:, None], m[j][:, None]
            v[i] -= mj / (mi + mj) * vr
            v[j] += mi / (mi + mj) * vr
    v -= (m[:, None] * v).sum(0) / m.sum()
    ndf = sys.ndf() - 3
    t_now = (m[:, None] * v * v).sum() * 1e7 / (GASCST * ndf)
    return v * math.sqrt(temp / t_now)


def _groups_by_size(sys: System):
    by = {}
    for atoms, _ in sys.shake_groups():
        by.setdefault(len(atoms), []).append(atoms)
    return [by[k] for k in sorted(by)]


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------

def water_box(nwat: int, L: float, seed: int, n_ion_pairs: int = 0, temp: float = 300.0,
              name: str = "water") -> System:
    """nwat rigid waters (+ optional Na/Cl pairs on vacant lattice sites) in
    a cubic periodic box [0, L)^3."""
    rng = np.random.Generator(np.random.PCG64(seed))
    Lv = np.array([L, L, L], np.float64)
    nion = 2 * n_ion_pairs
    sites = _lattice_sites(nwat + nion, L)
    pick = rng.permutation(len(sites))[: nwat + nion]
    sites = sites[np.sort(pick)]
    sites = sites + rng.uniform(-0.2, 0.2, size=sites.shape)
    ion_idx = np.sort(rng.permutation(len(sites))[:nion]) if nion else np.zeros(0, int)
    wmask = np.ones(len(sites), bool)
    wmask[ion_idx] = False
    ions = sites[ion_idx]
    pos_w = _place_waters(rng, sites[wmask], Lv, fixed=ions if nion else None)
    types = water_types()
    mols = [water_template(nwat)]
    x = [pos_w]
    if nion:
        types += [AtomType("NA", NA_RMIN2, NA_EPS), AtomType("CL", CL_RMIN2, CL_EPS)]
        mols += [ion_template("NA", n_ion_pairs, 3, NA_MASS, 1.0), ion_template("CL", n_ion_pairs, 4, CL_MASS, -1.0)]
        x += [ions[:n_ion_pairs], ions[n_ion_pairs:]]
    xx = np.mod(np.concatenate(x), L)
    s = System(name=name, mols=mols, types=types, box=Lv, x=xx, v=np.zeros_like(xx))
    s.v = thermal_velocities(rng, s, temp)
    return s


R_CH = 1.09
MASS_C = 12.011


def methyl_template(ncopy: int, type_c: int, type_h: int) -> MolTemplate:
    """A rigid CH3 (C + 3 H, charges -0.3 / +0.1): one 4-atom SHAKE group
    (md_run's shake4: pairs 1-2, 2-3, 1-3, 1-4, 3-4, 2-4) with tetrahedral
    H-C-H angles."""
    r_hh = R_CH * math.sqrt(8.0 / 3.0)
    return MolTemplate(
        name="MET", ncopy=ncopy, aname=["C", "H1", "H2", "H3"], atname=["CT", "HC", "HC", "HC"],
        itype=np.array([type_c, type_h, type_h, type_h]), rname=["MET"] * 4, resnum=np.array([1, 1, 1, 1]),
        mass=np.array([MASS_C, MASS_H, MASS_H, MASS_H]), charge=np.array([-0.3, 0.1, 0.1, 0.1]),
        n1x=np.array([[3, 0, 0], [0, 2, 0], [0, 1, 0], [0, 0, 0]]), partners=[[1, 2, 3], [1, 2], [1], []],
        bonds=np.array([[1, 2], [1, 3], [1, 4]]), bond_fc=np.full(3, 340.0), bond_r0=np.full(3, R_CH),
        angles=np.array([[2, 1, 3], [2, 1, 4], [3, 1, 4]]), angle_fc=np.full(3, 35.0),
        angle_deg=np.full(3, math.degrees(math.acos(-1.0 / 3.0))),
        shake=[((1, 2, 3, 4), (R_CH, r_hh, R_CH, R_CH, r_hh, r_hh))])


def methyl_water(nmet: int, nwat: int, L: float, seed: int, temp: float = 300.0,
                 name: str = "methyl_water") -> System:
    """nmet rigid CH3 groups (4-atom SHAKE groups) and nwat waters in a
    periodic box: a water_box of nmet + nwat waters whose first nmet waters
    become CH3 groups on their oxygen sites, randomly oriented."""
    w = water_box(nmet + nwat, L, seed, temp=temp)
    rng = np.random.Generator(np.random.PCG64(seed + 15485863))
    tet = np.array([[1.0, 1.0, 1.0], [1.0, -1.0, -1.0], [-1.0, 1.0, -1.0]]) / math.sqrt(3.0)
    xm = []
    for i in range(nmet):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        a, b, c, d = q
        rot = np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                        [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                        [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])
        o = w.x[3 * i]
        xm.append(np.vstack([o, o + R_CH * tet @ rot.T]))
    types = water_types() + [AtomType("CT", 1.908, 0.1094), AtomType("HC", 1.487, 0.0157)]
    x = np.mod(np.concatenate(xm + [w.x[3 * nmet:]]), L)
    s = System(name=name, mols=[methyl_template(nmet, 3, 4), water_template(nwat)], types=types,
               box=w.box.copy(), x=x, v=np.zeros_like(x))
    s.v = thermal_velocities(rng, s, temp)
    return s


def water_droplet(R: float, seed: int, density_L: float = 31.04, temp: float = 300.0,
                  name: str = "droplet") -> System:
    """A non-periodic water droplet for the cell multipole (CMM) path: the
    waters of a water_box at 0.997 g/cm^3 whose oxygen lies within R of the
    box centre.  The box only places the droplet (its centre is the box
    centre); the system is not periodic."""
    L = max(density_L, 2.0 * R + 8.0)
    nwat = int(round(L ** 3 * 0.997 / 18.015 * 0.6022))
    w = water_box(nwat, L, seed, temp=temp)
    c = 0.5 * L
    ox = w.x[0::3]
    keep = np.where(((ox - c) ** 2).sum(axis=1) <= R * R)[0]
    idx = (3 * keep[:, None] + np.arange(3)[None, :]).ravel()
    s = System(name=name, mols=[water_template(len(keep))], types=water_types(), box=w.box.copy(),
               x=w.x[idx].copy(), v=np.zeros((len(idx), 3)))
    rng = np.random.Generator(np.random.PCG64(seed + 7919))
    s.v = thermal_velocities(rng, s, temp)
    return s


def chain_droplet(n_chain: int, R: float, seed: int, lmd: Optional[float] = None, bond_fc: float = 0.0,
                  angle_fc: float = 0.0, torsion_fc: float = 0.0, temp: float = 300.0,
                  name: str = "chain_droplet") -> System:
    """A non-periodic droplet around a compact chain (cell multipole path):
    a chain_in_water system with a globular chain at the box centre, of which
    the chain and the waters whose oxygen lies within R of the centre are
    kept."""
    L = 2.0 * R + 10.0
    w = chain_in_water(n_chain, L, seed, lmd=lmd, bond_fc=bond_fc, angle_fc=angle_fc, torsion_fc=torsion_fc,
                       compact=True, temp=temp)
    chain, water = w.mols
    nc = chain.natom
    c = 0.5 * L
    ox = w.x[nc::3]
    keep = np.where(((ox - c) ** 2).sum(axis=1) <= R * R)[0]
    widx = nc + (3 * keep[:, None] + np.arange(3)[None, :]).ravel()
    idx = np.concatenate([np.arange(nc), widx])
    s = System(name=name, mols=[chain, water_template(len(keep), cluster=water.cluster)], types=w.types,
               box=w.box.copy(), x=w.x[idx].copy(), v=np.zeros((len(idx), 3)),
               cluster_method=w.cluster_method, lam=w.lam)
    rng = np.random.Generator(np.random.PCG64(seed + 104729))
    s.v = thermal_velocities(rng, s, temp)
    return s


def _saw_chain(rng, n: int, L: float, bond: float = 1.5, excl: float = 3.0,
               min_angle_deg: float = 100.0, confine: Optional[float] = None) -> np.ndarray:
    """Self-avoiding random walk in a periodic box: atom k keeps >= excl from
    every atom k' <= k-3 (minimum image) and bond angles >= min_angle.  With
    `confine`, every atom stays within that radius of the box centre (a
    compact, globule-like chain); the radius grows by 2% after every 20
    backtracks so that a trapped walk always gets out."""
    cell = excl
    nc = max(1, int(L // cell))
    cs = L / nc
    grid = {}
    pos = np.zeros((n, 3))
    pos[0] = L / 2

    def key(p):
        return tuple((np.floor(np.mod(p, L) / cs).astype(int) % nc).tolist())

    def ok(p, k):
        c = key(p)
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for j in grid.get(((c[0] + dx) % nc, (c[1] + dy) % nc, (c[2] + dz) % nc), ()):
                        if j >= k - 2:
                            continue
                        d = p - pos[j]
                        d -= L * np.round(d / L)
                        if d @ d < excl * excl:
                            return False
        return True

    cosmax = math.cos(math.radians(min_angle_deg))
    grid.setdefault(key(pos[0]), []).append(0)   # the first atom excludes like every other
    nback = 0
    k = 1
    while k < n:
        placed = False
        for _ in range(200):
            u = rng.normal(size=3)
            u /= np.linalg.norm(u)
            if k >= 2:
                prev = pos[k - 1] - pos[k - 2]
                prev /= np.linalg.norm(prev)
                if -(u @ prev) > cosmax:      # angle(k-2,k-1,k) must exceed min_angle
                    continue
            p = pos[k - 1] + bond * u
            if confine is not None and np.sum((p - L / 2) ** 2) > confine * confine:
                continue
            if ok(p, k):
                placed = True
                break
        if not placed:                         # backtrack a few atoms
            nback += 1
            if confine is not None and nback % 20 == 0:
                confine *= 1.02
            back = min(k - 1, 5)
            for j in range(k - back, k):
                grid[key(pos[j])].remove(j)
            k -= back
            continue
        pos[k] = p
        grid.setdefault(key(p), []).append(k)
        k += 1
    return pos


def chain_template(rng, n: int, type0: int, ntypes: int, cluster: int = 1,
                   bond_fc: float = 0.0, angle_fc: float = 0.0, pos: Optional[np.ndarray] = None,
                   L: float = 0.0, bond_r0: Optional[float] = None,
                   angle_deg: Optional[float] = None,
                   torsion_fc: float = 0.0) -> Tuple[MolTemplate, List[AtomType]]:
    """Linear chain: bonds i-(i+1), angles i-(i+1)-(i+2); exclusions are the
    1-2 and 1-3 pairs; every 4th atom (index 3, 7, ...) is 'H' constrained to
    its predecessor.  Harmonic bond/angle force constants default to 0 (no
    bonded forces); their equilibrium values are the generated geometry's own
    unless `bond_r0` / `angle_deg` fix them.  With torsion_fc > 0 every
    i..i+3 is a torsion (periodicity 1-3, phase 0 or 180 deg, force constant
    torsion_fc * U[0.5, 1.5], 1-4 flag set), and the 1-4 pairs join the
    exclusions."""
    q = rng.normal(0.0, 0.3, size=n)
    q -= q.mean()
    sig = rng.uniform(1.9, 3.8, size=ntypes)
    eps = rng.uniform(0.015, 0.2, size=ntypes)
    types = [AtomType(f"C{t:02d}", 2.0 ** (1.0 / 6.0) * sig[t] / 2.0, eps[t]) for t in range(ntypes)]
    itype = type0 + rng.integers(0, ntypes, size=n)
    is_h = (np.arange(n) % 4) == 3
    mass = np.where(is_h, MASS_H, 12.011)
    aname = [("H" if h else "C") + str(i + 1) for i, h in enumerate(is_h)]
    n1x = np.zeros((n, 3), int)
    partners = []
    for i in range(n):
        p = []
        if i + 1 < n:
            p.append(1)
            n1x[i, 0] = 1
        if i + 2 < n:
            p.append(2)
            n1x[i, 1] = 1
        if torsion_fc > 0 and i + 3 < n:
            p.append(3)
            n1x[i, 2] = 1
        partners.append(p)
    bonds = np.stack([np.arange(1, n), np.arange(2, n + 1)], axis=1)
    angles = np.stack([np.arange(1, n - 1), np.arange(2, n), np.arange(3, n + 1)], axis=1)
    if pos is not None:
        d = pos[1:] - pos[:-1]
        d -= L * np.round(d / L)
        r0 = np.linalg.norm(d, axis=1)
        a1 = -d[:-1]
        a2 = d[1:]
        cosang = (a1 * a2).sum(1) / (np.linalg.norm(a1, axis=1) * np.linalg.norm(a2, axis=1))
        th0 = np.degrees(np.arccos(np.clip(cosang, -1, 1)))
    else:
        r0 = np.full(n - 1, 1.5)
        th0 = np.full(n - 2, 109.5)
    if bond_r0 is not None:
        r0 = np.full(n - 1, float(bond_r0))
    if angle_deg is not None:
        th0 = np.full(n - 2, float(angle_deg))
    shake = [((i, i + 1), (1.5,)) for i in range(3, n + 1, 4) if i <= n and i >= 1 and i + 1 <= n]
    shake = [((a, b), d) for (a, b), d in shake if is_h[b - 1]]
    tors = dict()
    if torsion_fc > 0 and n >= 4:
        nt = n - 3
        tors = dict(torsions=np.stack([np.arange(1, nt + 1) + k for k in range(4)], axis=1),
                    tors_fc=torsion_fc * rng.uniform(0.5, 1.5, size=nt), tors_div=np.ones(nt, int),
                    tors_rot=rng.integers(1, 4, size=nt).astype(np.float64),
                    tors_phase_deg=np.where(rng.random(nt) < 0.5, 0.0, 180.0), tors_nbf=np.ones(nt, int))
    mol = MolTemplate(
        name="CHAIN", ncopy=1, aname=aname, atname=["HC" if h else "CT" for h in is_h],
        itype=itype, rname=["RES"] * n, resnum=np.arange(1, n + 1), mass=mass, charge=q,
        n1x=n1x, partners=partners, bonds=bonds, bond_fc=np.full(n - 1, bond_fc), bond_r0=r0,
        angles=angles, angle_fc=np.full(n - 2, angle_fc), angle_deg=th0, shake=shake, cluster=cluster, **tors)
    return mol, types


def chain_in_water(n_chain: int, L: float, seed: int, n_chain_types: int = 16, temp: float = 300.0,
                   lmd: Optional[float] = None, name: str = "chain", bond_fc: float = 0.0,
                   angle_fc: float = 0.0, bond_r0: Optional[float] = None,
                   angle_deg: Optional[float] = None, compact: bool = False, torsion_fc: float = 0.0,
                   solvent: str = "lattice", n_target: Optional[int] = None) -> System:
    """Solvated chain: SAW chain in a periodic water box.  With `lmd`, the
    chain is cluster class 1 and water class 2 under CLSTMETHOD=LMD (three
    energy classes: chain-chain lambda^2, chain-water lambda, water-water 1).

    solvent="lattice": every site of the smallest cubic lattice with at least
    0.997 g/cm^3 worth of sites; waters with any atom within 2.6 A of a chain
    atom are removed (O sites within 3.96 A never get a water).
    solvent="bulk": the water of the 6k-atom water box (2,000 waters in a
    39.1 A box, 1.0007 g/cm^3) on a random subset of lattice sites; a water
    is removed when its O is within 2.5 A or an H within 1.8 A of a chain
    atom (O sites within 2.8 A never get a water).  With n_target, waters
    drawn at random are removed until the system has n_target atoms (or the
    nearest count above it that whole waters allow)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    Lv = np.array([L, L, L], np.float64)
    if compact:
        # globule of ~23 A^3 per chain atom to start with (non-neighbours >= 3.0 A apart)
        cpos = _saw_chain(rng, n_chain, L, excl=3.0, confine=(n_chain * 23.0 * 3 / (4 * math.pi)) ** (1 / 3))
    else:
        cpos = _saw_chain(rng, n_chain, L)
    # bond lengths of the walk are exactly 1.5 A, so SHAKE targets are met
    from scipy.spatial import cKDTree
    ctree = cKDTree(np.mod(cpos, L), boxsize=Lv)
    if solvent == "lattice":
        nsite = int(round(WATER_PER_A3 * L ** 3))
        sites = _lattice_sites(nsite, L)
        pre, o_gap, h_gap = 2.6 + R_OH + 0.4, 2.6, 2.6
    elif solvent == "bulk":
        nsite = int(round(BULK_WATER_PER_A3 * L ** 3))
        sites = _lattice_sites(nsite, L)
        sites = sites[np.sort(rng.permutation(len(sites))[:nsite])]
        pre, o_gap, h_gap = 2.8, 2.5, 1.8
    else:
        raise ValueError(f"unknown solvent mode {solvent!r}")
    sites = sites + rng.uniform(-0.2, 0.2, size=sites.shape)
    near = ctree.query_ball_point(np.mod(sites, L), r=pre)
    keep = np.array([len(nb) == 0 for nb in near])
    pos_w = _place_waters(rng, sites[keep], Lv, fixed=cpos)
    d, _ = ctree.query(np.mod(pos_w, L))
    d = d.reshape(-1, 3)
    bad = (d[:, 0] < o_gap) | (d[:, 1:] < h_gap).any(1)
    pos_w = pos_w.reshape(-1, 3, 3)[~bad]
    if n_target is not None:
        nw_want = max(0, -(-(int(n_target) - n_chain) // 3))
        if len(pos_w) > nw_want:
            pos_w = pos_w[np.sort(rng.permutation(len(pos_w))[:nw_want])]
    pos_w = pos_w.reshape(-1, 3)
    nwat = len(pos_w) // 3
    lm = lmd is not None
    wt = water_types()
    chain, ctypes = chain_template(rng, n_chain, type0=len(wt) + 1, ntypes=n_chain_types,
                                   cluster=1, bond_fc=bond_fc, angle_fc=angle_fc, pos=cpos, L=L,
                                   bond_r0=bond_r0, angle_deg=angle_deg, torsion_fc=torsion_fc)
    mols = [chain, water_template(nwat, cluster=2 if lm else 1)]
    xx = np.mod(np.concatenate([cpos, pos_w]), L)
    s = System(name=name, mols=mols, types=wt + ctypes, box=Lv, x=xx, v=np.zeros_like(xx),
               cluster_method="LMD" if lm else "AND", lam=float(lmd) if lm else 0.0)
    s.v = thermal_velocities(rng, s, temp)
    return s


# ---------------------------------------------------------------------------
# the benchmark workloads
# ---------------------------------------------------------------------------

# harmonic constants of the benchmark chains (kcal/mol/A^2, kcal/mol/rad^2),
# AMBER-like C-C values; equilibrium values are the generated geometry's
CHAIN_BOND_FC, CHAIN_ANGLE_FC = 300.0, 50.0

WORKLOADS = {
    0: "water-6k: 2,000 TIP3P waters, 39.1 A box",
    1: "chain-22k: 2,500-atom chain (harmonic bonds/angles) in water, 64 A box (22,297 atoms)",
    2: "water-100k: 33,333 waters + 60 Na/Cl pairs, 100.0 A box",
    3: "ensemble-64: 64 chain-in-water systems (harmonic bonds/angles), 3k-30k atoms each",
    4: "water-1M: 333,333 waters, 216 A box",
}


def workload(i: int) -> System:
    seed = 1000 + i
    if i == 0:
        return water_box(2000, 39.1, seed, name="water-6k")
    if i == 1:
        return chain_in_water(2500, 64.0, seed, name="chain-22k", bond_fc=CHAIN_BOND_FC, angle_fc=CHAIN_ANGLE_FC,
                              compact=True, solvent="bulk")
    if i == 2:
        L = (33333 / WATER_PER_A3) ** (1.0 / 3.0)
        return water_box(33333, round(L, 3), seed, n_ion_pairs=60, name="water-100k")
    if i == 4:
        return water_box(333333, 216.0, seed, name="water-1M")
    raise ValueError("workload 3 is an ensemble; use ensemble()")


def ensemble_sizes(nsys: int = 64, seed: int = 1003, lo: int = 3000, hi: int = 30000):
    """Target atom counts (log-uniform in [lo, hi]) and lambdas (U[0.5, 1.0])
    of the ensemble members."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = np.round(np.exp(rng.uniform(math.log(lo), math.log(hi), size=nsys))).astype(int)
    lams = rng.uniform(0.5, 1.0, size=nsys)
    return sizes, lams


def _ensemble_member(args) -> System:
    k, ns, lam, seed = args
    nch = max(8, int(round(0.104 * ns)) // 4 * 4)
    # box for ns atoms: the waters at bulk density plus the chain's displaced volume
    # (slightly overestimated; the surplus waters are removed at random)
    L = ((ns - nch) / (3 * BULK_WATER_PER_A3) + nch * CHAIN_ATOM_A3) ** (1.0 / 3.0)
    L = max(L, 24.5)
    return chain_in_water(nch, round(L, 3), seed * 1000 + k, lmd=float(lam), name=f"ens{k:02d}",
                          bond_fc=CHAIN_BOND_FC, angle_fc=CHAIN_ANGLE_FC, compact=True,
                          solvent="bulk", n_target=int(ns))


def ensemble(nsys: int = 64, seed: int = 1003, lo: int = 3000, hi: int = 30000, members=None,
             workers: int = 8) -> List[System]:
    """Uneven ensemble: sizes log-uniform in [lo, hi] atoms; each member a
    chain-in-water system (chain = 10.4% of atoms) with its own lambda in
    [0.5, 1.0] on the chain's interactions.  `members`: indices of the members
    to build (default all).  Members are independent (own seeds), so they are
    built in `workers` processes with the same result as serially."""
    sizes, lams = ensemble_sizes(nsys, seed, lo, hi)
    idx = list(range(nsys)) if members is None else list(members)
    jobs = [(k, int(sizes[k]), float(lams[k]), seed) for k in idx]
    if workers <= 1 or len(jobs) <= 1:
        return [_ensemble_member(j) for j in jobs]
    import concurrent.futures as cf
    import multiprocessing as mp
    with cf.ProcessPoolExecutor(max_workers=min(workers, len(jobs)), mp_context=mp.get_context("fork")) as ex:
        return list(ex.map(_ensemble_member, jobs))
