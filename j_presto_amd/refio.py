"""Readers and writers for md_run's input and output files.

Formats (reference file:line):
  .tpl  free-format topology, `;` comments, `->` continuation
        (src_md_run/inptpl.f90:62-140 reader; rdatom :1051-1134, rdbond :1140,
        rdangl :1192, rdfunc :1360, rdnonb/rdnoam :1427-1571; sp/prep:762-841
        is the reference writer, header line at sp/prep:1040)
  .shk  `SHAKE>` blocks: molecule name, then `natom a1..an d12 d23 d13 ...`
        (setshk.f90:120-400; pair order 1-2, 2-3, 1-3, 1-4, 3-4, 2-4 from
        parlis at setshk.f90:541-542 and mkshkl.f90:674-690)
  .crd  binary coordinates, CRDFMT=BINA (inpcrb.f90:30-40)
  .pdb  CRDFMT=PDB: ATOM records, x y z as 3f8.3 after column 30 (inppdb.f90)
  restart  Fortran unformatted: title; natm,nvar; loop,time,Et,Ek,Ep;
        cord(3,N); vel(3,nvar); lambda,lambda_v (outrst.f90:56-61)
  1stp_force.bin  FDEBUG=MAKE dump: epot(15); lambda_f; grad(3,N)
        (verlethe.f90:378-384)
All Fortran unformatted files use 4-byte record markers (gfortran/flang
default on x86-64).
"""
from __future__ import annotations

import os
import struct
from typing import Dict, List, Optional, Sequence

import numpy as np

from .systems import System


# --------------------------------------------------------------------------
# Fortran sequential unformatted records
# --------------------------------------------------------------------------

def write_records(path: str, records: Sequence[bytes]) -> None:
    with open(path, "wb") as f:
        for r in records:
            f.write(struct.pack("<i", len(r)))
            f.write(r)
            f.write(struct.pack("<i", len(r)))


def read_records(path: str) -> List[bytes]:
    out = []
    with open(path, "rb") as f:
        data = f.read()
    p = 0
    while p < len(data):
        (n,) = struct.unpack_from("<i", data, p)
        out.append(data[p + 4:p + 4 + n])
        (m,) = struct.unpack_from("<i", data, p + 4 + n)
        if m != n:
            raise ValueError(f"{path}: corrupt record marker at byte {p}")
        p += 8 + n
    return out


# --------------------------------------------------------------------------
# restart / coordinates / force dump
# --------------------------------------------------------------------------

def write_restart(path: str, x: np.ndarray, v: np.ndarray, loop: int = 0, time_fs: float = 0.0,
                  et: float = 0.0, ek: float = 0.0, ep: float = 0.0, lam: float = 0.0,
                  lam_v: float = 0.0) -> None:
    n = len(x)
    title = b"     VERLET synthetic".ljust(80)
    write_records(path, [
        title,
        struct.pack("<ii", n, len(v)),
        struct.pack("<idddd", loop, time_fs, et, ek, ep),
        np.ascontiguousarray(x, np.float64).tobytes(),
        np.ascontiguousarray(v, np.float64).tobytes(),
        struct.pack("<dd", lam, lam_v),
    ])


def read_restart(path: str) -> Dict[str, object]:
    r = read_records(path)
    n, nvar = struct.unpack("<ii", r[1])
    loop, t, et, ek, ep = struct.unpack("<idddd", r[2])
    x = np.frombuffer(r[3], np.float64).reshape(n, 3).copy()
    v = np.frombuffer(r[4], np.float64).reshape(nvar, 3).copy()
    lam, lam_v = struct.unpack("<dd", r[5])
    return dict(title=r[0].decode(errors="replace"), n=n, nvar=nvar, loop=loop, time=t,
                et=et, ek=ek, ep=ep, x=x, v=v, lam=lam, lam_v=lam_v)


def write_crd(path: str, x: np.ndarray) -> None:
    x = np.asarray(x, np.float64)
    write_records(path, [
        b"synthetic".ljust(23) + b"j_presto_amd".ljust(40),
        struct.pack("<i", len(x)),
        x[:, 0].tobytes(), x[:, 1].tobytes(), x[:, 2].tobytes(),
    ])


def write_pdb(path: str, x: np.ndarray) -> None:
    """PDB coordinates as md_run reads them (inppdb.f90: ATOM records, x y z
    in columns 31-54 as 3f8.3)."""
    x = np.asarray(x, np.float64)
    L = []
    for i, (a, b, c) in enumerate(x):
        L.append(f"ATOM  {(i + 1) % 100000:>5d} X    UNK A{(i // 3 + 1) % 10000:>4d}    {a:8.3f}{b:8.3f}{c:8.3f}"
                 f"{1.0:6.2f}{0.0:6.2f}")
    L.append("END")
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def read_force_dump(path: str, n: int) -> Dict[str, np.ndarray]:
    r = read_records(path)
    epot = np.frombuffer(r[0], np.float64).copy()
    (lam_f,) = struct.unpack("<d", r[1])
    grad = np.frombuffer(r[2], np.float64).reshape(n, 3).copy()
    return dict(epot=epot, lambda_f=lam_f, grad=grad)


# --------------------------------------------------------------------------
# topology and SHAKE files
# --------------------------------------------------------------------------

def _fmt(x: float) -> str:
    return repr(float(x))


def _pack_tokens(toks: Sequence[str], width: int = 72) -> List[str]:
    """One free-format record split over lines of < 80 columns (md_run
    reads lines as a80) joined by `->` continuations."""
    lines, cur = [], ""
    for t in toks:
        if cur and len(cur) + 1 + len(t) > width:
            lines.append(cur + " ->")
            cur = t
        else:
            cur = t if not cur else cur + " " + t
    lines.append(cur)
    return lines


def write_tpl(path: str, s: System) -> None:
    maxatm = s.n
    maxbnd = sum(len(m.bonds) * m.ncopy for m in s.mols)
    maxang = sum(len(m.angles) * m.ncopy for m in s.mols)
    maxtor = sum(len(m.torsions) * m.ncopy for m in s.mols)
    L = [f"; {maxatm} {max(maxbnd, 1)} {max(maxang, 1)} {max(maxtor, 1)} 1 {len(s.types)}", "",
         "TPL> TITLE", f"{s.name}", "", "TPL> MOLECULES"]
    for m in s.mols:
        L.append(f"{m.name:<40}{m.ncopy:>12}")
    for m in s.mols:
        L += ["", "TPL> ATOMS", m.name]
        for a in range(m.natom):
            n1x = m.n1x[a]
            rmin2 = s.types[int(m.itype[a]) - 1].rmin2
            toks = [m.aname[a], m.atname[a], str(int(m.itype[a])), m.rname[a], str(int(m.resnum[a])),
                    _fmt(m.mass[a]), _fmt(rmin2), _fmt(m.charge[a]), str(n1x[0]), str(n1x[1]), str(n1x[2])]
            toks += [str(p) for p in m.partners[a]]
            toks += ["0", "0", "0", "0", "0.0", "0.0", "0.0"]
            L += _pack_tokens(toks)
    for m in s.mols:
        if len(m.bonds):
            L += ["", "TPL> BONDS", m.name]
            for (i, j), fc, r0 in zip(m.bonds, m.bond_fc, m.bond_r0):
                L.append(f"{i} {j} {_fmt(fc)} {_fmt(r0)}")
    for m in s.mols:
        if len(m.angles):
            L += ["", "TPL> ANGLES", m.name]
            for (i, j, k), fc, th in zip(m.angles, m.angle_fc, m.angle_deg):
                L.append(f"{i} {j} {k} {_fmt(fc)} {_fmt(th)}")
    for m in s.mols:
        if len(m.torsions):   # inptpl.f90 rdtors: i j k l fc div rot phase(deg) 1-4 flag
            L += ["", "TPL> TORSIONS", m.name]
            for (i, j, k, l), fc, dv, rot, ph, nbf in zip(m.torsions, m.tors_fc, m.tors_div, m.tors_rot,
                                                          m.tors_phase_deg, m.tors_nbf):
                L.append(f"{i} {j} {k} {l} {_fmt(fc)} {int(dv)} {_fmt(rot)} {_fmt(ph)} {int(nbf)}")
    L += ["", "TPL> FUNCTIONS", "1 4 LENNARD-JONES-AMBER", "", "TPL> NONBONDS"]
    for t, at in enumerate(s.types):
        # inptpl.f90 NONB record: type, ..., Rmin/2, eps, 1-4 elec scale (fy14se), 1-4 vdW scale (fy14sv)
        L.append(f"{t + 1} 0 1 {_fmt(at.rmin2)} {_fmt(at.eps)} {_fmt(at.scale14_elec)} {_fmt(at.scale14_vdw)}"
                 f" ; {at.name}")
    L.append("")
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def write_shk(path: str, s: System) -> None:
    L = []
    for m in s.mols:
        if not m.shake:
            continue
        L += ["", "SHAKE> SHAKE", m.name]
        for atoms, dist in m.shake:
            L.append(" ".join([str(len(atoms))] + [str(a) for a in atoms] + [_fmt(d) for d in dist]))
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def write_md_inp(path: str, files: Dict[str, str], s: System, md: Dict[str, str],
                 inp: Dict[str, str] = None) -> None:
    """md.inp with an INPUT and an MD phase.  `files` has TOPOLO, COORDI
    (binary) and optionally SETSHK; `md` is extra EXE> MD keywords, `inp`
    extra EXE> INPUT keywords (e.g. LWEIGH, LTCNTL of inpinp.f90)."""
    L = ["EXE> INPUT", f"      TOPOLO= {files['TOPOLO']}", f"      COORDI= {files['COORDI']}",
         "      CRDFMT= BINA"]
    if files.get("SETSHK"):
        L.append(f"      SETSHK= {files['SETSHK']}")
    if s.cluster_method == "LMD":
        na = int((s.cluster == 1).sum())
        L.append(f"      CLUSTA= 1<{na}")
        L.append("      GEPSME= LMD")
    for k, val in (inp or {}).items():
        L.append(f"      {k}= {val}")
    L += ["QUIT", "", "EXE> MD"]
    box = s.celwal
    base = {"CAL15M": "ZD", "CUTLEN": _fmt(s.cutoff), "BOUNDA": "PBC",
            "XMINCL": _fmt(box[0]), "XMAXCL": _fmt(box[1]), "YMINCL": _fmt(box[2]),
            "YMAXCL": _fmt(box[3]), "ZMINCL": _fmt(box[4]), "ZMAXCL": _fmt(box[5]),
            "PARALV": "DOUB", "N_SIMD": "1", "STOPCE": "NO"}
    base.update(md)
    for k, val in base.items():
        L.append(f"      {k}= {val}")
    L += ["QUIT", "", "EXE> END", ""]
    with open(path, "w") as f:
        f.write("\n".join(L))


def write_inputs(dirname: str, s: System, shake: bool = True) -> Dict[str, str]:
    os.makedirs(dirname, exist_ok=True)
    files = {"TOPOLO": "sys.tpl", "COORDI": "sys.crd"}
    write_tpl(os.path.join(dirname, "sys.tpl"), s)
    write_crd(os.path.join(dirname, "sys.crd"), s.x)
    if shake and any(m.shake for m in s.mols):
        write_shk(os.path.join(dirname, "sys.shk"), s)
        files["SETSHK"] = "sys.shk"
    return files


# --------------------------------------------------------------------------
# system file of the Fortran host program (fortran/jp_host.f90)
# --------------------------------------------------------------------------

def write_jpsys(path: str, s: System, dt: float = 2.0, cell_level: int = 3, shake_tol: float = 1e-6,
                shake_maxit: int = 1000) -> None:
    """The descriptor of `s` (the arrays jp_sys_desc points to) plus its
    velocities, one array per Fortran record, in the order jp_host reads."""
    from .abi import DescHolder
    h = DescHolder(s, dt=dt, cell_level=cell_level, shake_tol=shake_tol, shake_maxit=shake_maxit)
    d = h.desc
    ng = d.nshake
    hdr = np.array([d.natom, d.ntyp, d.nfrag, d.nfrg, d.nexcl, d.nexcess, ng, shake_maxit, cell_level, d.lmd,
                    d.nbond, d.nangle, d.ntors, d.n14], np.int32)
    rhdr = np.array([shake_tol, d.cutoff, d.diel, dt, d.lambda_] + [float(w) for w in d.celwal], np.float64)
    i32 = lambda a: np.ascontiguousarray(a, np.int32).tobytes()
    f64 = lambda a: np.ascontiguousarray(a, np.float64).tobytes()
    write_records(path, [
        hdr.tobytes(), rhdr.tobytes(), f64(h.x), f64(s.v), f64(h.mass), f64(h.charge), i32(h.atype), i32(h.icls),
        f64(h.rmin2), f64(h.eps), i32(h.nmat), i32(h.excl), i32(h.excess), i32(h.shk_n[:ng]), i32(h.shk_a[:ng]),
        f64(h.shk_d[:ng]), i32(h.bond_atoms), f64(h.bond_fc), f64(h.bond_r0), i32(h.angle_atoms),
        f64(h.angle_fc), f64(h.angle_theta0), i32(h.tors_atoms), f64(h.tors_fc), f64(h.tors_rot),
        f64(h.tors_phase), i32(h.tors_div), i32(h.p14_atoms), f64(h.p14_vdw), f64(h.p14_elec),
    ])


def read_jphost(path: str, n: int, nfrg: int) -> Dict[str, object]:
    """Output file of jp_host."""
    r = read_records(path)
    f = lambda b: np.frombuffer(b, np.float64)
    out = {
        "ier": np.frombuffer(r[0], np.int32).copy(),
        "grad64": f(r[1]).reshape(nfrg, n, 3).copy(), "eelec64": f(r[2]).copy(), "evdw64": f(r[3]).copy(),
        "grad32": f(r[4]).reshape(nfrg, n, 3).copy(), "eelec32": f(r[5]).copy(), "evdw32": f(r[6]).copy(),
        "eterm": f(r[7]).reshape(nfrg, 6).copy(),
        "x": f(r[8]).reshape(n, 3).copy(), "v": f(r[9]).reshape(n, 3).copy(),
    }
    nlog = struct.unpack("<i", r[10])[0]
    out["log"] = [struct.unpack("<iddd", r[11 + k]) for k in range(nlog)]
    return out
