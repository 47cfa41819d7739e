// Cell multipole method for non-periodic systems (md_run CAL15M= CMM with
// BOUNDA= NO, PARALV= DOUB): FP64 throughout, as md_run.
//
// Cells: nlev levels, nv(l) = 4 * 2^(nlev - 1 - l) cells per edge of size
// siz(l) = cminsiz * 2^l (0-based l; incel.f90:29-57), centred on the mean of
// the coordinates at every rebuild (celpar_CB, celpar.f90:49-58).
//
// Layout in HBM, rebuilt every UPDATE steps (celprp):
//   key of an atom = (cell * nf + fragment) * 2 + has-vdW, level-1 cell;
//   atoms sorted by key, and within a key by atom number: every level-1
//   cell is one contiguous range, EL-only atoms before the vdW ones per
//   fragment (md_run's ipcl order after celprp's EL split, celprp.f90:47-72);
//   sx[p] = (x, y, z, chgmod) of the sorted atom p, refreshed every step;
//   sx0[p] its coordinates at the rebuild (md_run's Tcord: the second-nearest
//   LJ table is cut at rlim with them, neartb.f90:321-366);
//   npc, pot: atoms and the 13 multipole coefficients of every (cell,
//   fragment) at every level (celoth, celpol).
// Per step:
//   k_cmm_gather   sx from x;
//   k_cmm_near     Coulomb of every pair in a cell and its 26 neighbours,
//                  LJ where both atoms have LJ (enear_CMM_doub's itbEL / itb
//                  tables), one wavefront per cell, lanes = (replica, i atom),
//                  replicas take every R-th j and are summed in a fixed order;
//   k_cmm_lj2      LJ of the second-nearest cells (Chebyshev distance 2 ..
//                  idelay) for vdW pairs within rlim at the rebuild;
//   k_cmm_far      when the far field is due (itwin): the particle-multipole
//                  field of every level's far cells (cmmdo.f90:164-288), one
//                  wavefront per cell, the multipoles wave-uniform;
//   k_cmm_combine  grad = sum over classes of factor * (near + far), into
//                  the slot gradient the integrators read (slot = atom).
// Excluded pairs (itab14) are skipped where md_run deletes them from the
// tables; a pair is tested only when both atoms lie in one molecule
// (connected component of the excluded-pair graph).
#pragma once
#include "jp_kernels.hpp"

#define JP_CMM_MAXLEV 5

struct CmmArgs {
    int n, nf, nfrg, nlev, idelay, nv1, nc1, ntyp;
    int nv[JP_CMM_MAXLEV];
    int cgoff[JP_CMM_MAXLEV];   // offset of level l's centres in cg
    int pcoff[JP_CMM_MAXLEV];   // offset of level l's (cell, fragment) entries in npc (x 13 in pot)
    double siz[JP_CMM_MAXLEV];
    double rlim2, ccoef;
    int ipmax, capel, capvd, capvw;   // ipmax, n15mxel, n15mx, nvdw (incel.f90:137-150, DOUB)
    int nkey;
    const double *cg;           // cell centres before the shift, per level
    double *c0;                 // [3] mean of the coordinates at the rebuild
    const double *x;            // coordinates, atom order
    const double *chg;          // chgmod = sqrt(ccoef) q
    const double *q;            // raw charges (celpol, chgmod2)
    const int *tflag;           // bits 0-11 type, 12 has-vdW, 13 cluster class
    const int *mol, *xptr, *xidx;
    const double2 *lj;          // (C6, C12) per type pair
    int *key, *kcnt, *kstart, *kfill, *tmp, *sorig, *spos, *sflag, *smol;
    double4 *sx;
    double *sx0;                // [n][3]
    int *npc;
    double *pot;
    int *err;                   // [0] bit 0 ier 28, bit 1 ier 51; [1] excluded pairs found (both ways)
    const SysP *sp;
    const double *efac;         // class factors of the system
    double *gnb;                // [p][3 classes][3] near-field sums of the sorted atom p
    double *enb;                // [p][6] near-field elec, vdW per class
    double *gwk;                // [a][3 classes][3] far-field gradient (kept between far-field steps)
    double *efar;               // [a][3] far-field energy per class (kept likewise)
    double *gslot;              // [a][3] combined gradient
    double *eatom;              // [a][6] elec (near + far), vdW per class; summed and halved by the host
};

__device__ __forceinline__ int cmm_lin(int nv, int x, int y, int z) { return (z * nv + y) * nv + x; }

// Energy class of a pair of cluster classes (n_matrix - 1).
__device__ __forceinline__ int cmm_cls(const SysP &P, int ci, int cj) { return P.cls[ci * 2 + cj]; }

// ------------------------------------------------------------------ rebuild

// Mean of the coordinates, fixed-order tree (celpar.f90:49-58 sums them and
// multiplies by 1/n).
__global__ void __launch_bounds__(1024) k_cmm_mean(CmmArgs A) {
#pragma clang fp contract(off)
    __shared__ double s[3][1024];
    const int t = threadIdx.x;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int i = t; i < A.n; i += 1024) {
        a0 = a0 + A.x[3 * i];
        a1 = a1 + A.x[3 * i + 1];
        a2 = a2 + A.x[3 * i + 2];
    }
    s[0][t] = a0; s[1][t] = a1; s[2][t] = a2;
    __syncthreads();
    for (int h = 512; h > 0; h >>= 1) {
        if (t < h)
            for (int k = 0; k < 3; ++k) s[k][t] = s[k][t] + s[k][t + h];
        __syncthreads();
    }
    if (t < 3) A.c0[t] = s[t][0] * (1.0 / (double)A.n);
}

__global__ void k_cmm_reset(CmmArgs A) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < A.nkey) { A.kcnt[i] = 0; A.kfill[i] = 0; }
    if (i < 2) A.err[i] = 0;
}

// Level-1 cell of every atom (celpar.f90:110-125): cells start half a cell
// below the first centre; int() truncates toward zero, so an atom is on the
// field when -1 < (x - origin) / siz < nv.  Off the field: ier 28
// (celpar.f90:184-192).
__global__ void k_cmm_bin(CmmArgs A) {
#pragma clang fp contract(off)
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= A.n) return;
    const double zz0 = A.siz[0] * 0.5, rr = 1.0 / A.siz[0];
    const double g0 = A.cg[A.cgoff[0]];
    int m[3];
    bool ok = true;
    for (int k = 0; k < 3; ++k) {
        const double o = (g0 + A.c0[k]) - zz0;
        const double f = (A.x[3 * a + k] - o) * rr;
        ok = ok && f > -1.0 && f < (double)A.nv1;
        m[k] = ok ? (int)f : 0;
    }
    if (!ok) {
        A.key[a] = -1;
        atomicOr(A.err, 1);
        return;
    }
    const int tf = A.tflag[a];
    const int cell = cmm_lin(A.nv1, m[0], m[1], m[2]);
    const int k = (cell * A.nf + ((tf >> 13) & 1)) * 2 + ((tf >> 12) & 1);
    A.key[a] = k;
    atomicAdd(A.kcnt + k, 1);
}

__global__ void k_cmm_place(CmmArgs A) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= A.n || (A.err[0] & 1)) return;
    const int k = A.key[a];
    A.tmp[A.kstart[k] + atomicAdd(A.kfill + k, 1)] = a;
}

// Rank within the key by atom number: the sorted order is independent of
// the arrival order of the atomics.
__global__ void k_cmm_rank(CmmArgs A) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= A.n || (A.err[0] & 1)) return;
    const int k = A.key[a], b = A.kstart[k], e = b + A.kcnt[k];
    int r = 0;
    for (int t = b; t < e; ++t) r += A.tmp[t] < a;
    A.sorig[b + r] = a;
    A.spos[a] = b + r;
}

__global__ void k_cmm_fill(CmmArgs A) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= A.n || (A.err[0] & 1)) return;
    const int a = A.sorig[p];
    const double x0 = A.x[3 * a], x1 = A.x[3 * a + 1], x2 = A.x[3 * a + 2];
    A.sx[p] = make_double4(x0, x1, x2, A.chg[a]);
    A.sx0[3 * p] = x0; A.sx0[3 * p + 1] = x1; A.sx0[3 * p + 2] = x2;
    A.sflag[p] = A.tflag[a];
    A.smol[p] = A.mol[a];
}

// celoth: atoms per (cell, fragment), level 1 from the keys (with celpar's
// ipmax test, celpar.f90:172-181), higher levels from the 8 children.
__global__ void k_cmm_npc(CmmArgs A, int l) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int mm = A.nv[l], nq = mm * mm * mm * A.nf;
    if (q >= nq || (A.err[0] & 1)) return;
    if (l == 0) {
        const int c = A.kcnt[2 * q] + A.kcnt[2 * q + 1];
        A.npc[q] = c;
        if (c > A.ipmax) atomicOr(A.err, 1);
        return;
    }
    const int f = q % A.nf, cell = q / A.nf;
    const int ix = cell % mm, iy = (cell / mm) % mm, iz = cell / (mm * mm), mc = A.nv[l - 1];
    int s = 0;
    for (int cz = 2 * iz; cz <= 2 * iz + 1; ++cz)
        for (int cy = 2 * iy; cy <= 2 * iy + 1; ++cy)
            for (int cx = 2 * ix; cx <= 2 * ix + 1; ++cx)
                s += A.npc[A.pcoff[l - 1] + cmm_lin(mc, cx, cy, cz) * A.nf + f];
    A.npc[A.pcoff[l] + q] = s;
}

// celpol: monopole, dipole and quadrupole of every (cell, fragment).  Level 1
// from the atoms (raw charges, about the shifted cell centre, in md_run's
// ipcl order), higher levels from the 8 children (celpol.f90:31-153).
__global__ void k_cmm_pol(CmmArgs A, int l) {
#pragma clang fp contract(off)
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int mm = A.nv[l], nq = mm * mm * mm * A.nf;
    if (q >= nq || (A.err[0] & 1)) return;
    const int f = q % A.nf, cell = q / A.nf;
    const int ix = cell % mm, iy = (cell / mm) % mm, iz = cell / (mm * mm);
    double *P = A.pot + (size_t)(A.pcoff[l] + q) * 13;
    if (A.npc[A.pcoff[l] + q] == 0) {
        for (int k = 0; k < 13; ++k) P[k] = 0.0;
        return;
    }
    const double *cgl = A.cg + A.cgoff[l];
    const double gx = cgl[ix] + A.c0[0], gy = cgl[iy] + A.c0[1], gz = cgl[iz] + A.c0[2];
    double qz = 0.0, rx = 0.0, ry = 0.0, rz = 0.0, rxx = 0.0, ryy = 0.0, rzz = 0.0, rxy = 0.0, ryz = 0.0, rzx = 0.0;
    if (l == 0) {
        const int b = A.kstart[2 * q], e = A.kstart[2 * q + 2];
        for (int p = b; p < e; ++p) {
            const int a = A.sorig[p];
            const double cq = A.q[a];
            qz = qz + cq;
            const double dx = A.sx0[3 * p] - gx, dy = A.sx0[3 * p + 1] - gy, dz = A.sx0[3 * p + 2] - gz;
            rx = rx + cq * dx; ry = ry + cq * dy; rz = rz + cq * dz;
            const double dxx = dx * dx, dyy = dy * dy, dzz = dz * dz;
            rxx = rxx + cq * (2.0 * dxx - dyy - dzz);
            ryy = ryy + cq * (2.0 * dyy - dzz - dxx);
            rzz = rzz + cq * (2.0 * dzz - dxx - dyy);
            rxy = rxy + cq * dx * dy;
            ryz = ryz + cq * dy * dz;
            rzx = rzx + cq * dz * dx;
        }
        const double v[13] = {qz, rx, ry, rz, rxx * 0.5, ryy * 0.5, rzz * 0.5, rxx, ryy, rzz,
                              rxy * 3.0, ryz * 3.0, rzx * 3.0};
        for (int k = 0; k < 13; ++k) P[k] = v[k];
        return;
    }
    const int mc = A.nv[l - 1];
    const double *cgc = A.cg + A.cgoff[l - 1];
    for (int jj = 0; jj < 8; ++jj) {
        const int mx = 2 * ix + (jj & 1), my = 2 * iy + ((jj >> 1) & 1), mz = 2 * iz + ((jj >> 2) & 1);
        const int qc = A.pcoff[l - 1] + cmm_lin(mc, mx, my, mz) * A.nf + f;
        if (A.npc[qc] == 0) continue;
        const double *C = A.pot + (size_t)qc * 13;
        const double qq = C[0];
        qz = qz + qq;
        const double xx = (cgc[mx] + A.c0[0]) - gx;
        const double yy = (cgc[my] + A.c0[1]) - gy;
        const double zz = (cgc[mz] + A.c0[2]) - gz;
        const double ax = C[1], ay = C[2], az = C[3];
        rx = rx + ax + qq * xx;
        ry = ry + ay + qq * yy;
        rz = rz + az + qq * zz;
        const double dxx = xx * xx, dyy = yy * yy, dzz = zz * zz;
        const double xax = xx * ax, yay = yy * ay, zaz = zz * az;
        rxx = rxx + C[4] + 2.0 * xax - yay - zaz + qq * (dxx - 0.5 * (dyy + dzz));
        ryy = ryy + C[5] + 2.0 * yay - zaz - xax + qq * (dyy - 0.5 * (dzz + dxx));
        rzz = rzz + C[6] + 2.0 * zaz - xax - yay + qq * (dzz - 0.5 * (dxx + dyy));
        rxy = rxy + C[10] * 0.5 + (yy * ax + xx * ay + xx * yy * qq) * 1.5;
        ryz = ryz + C[11] * 0.5 + (zz * ay + yy * az + yy * zz * qq) * 1.5;
        rzx = rzx + C[12] * 0.5 + (xx * az + zz * ax + zz * xx * qq) * 1.5;
    }
    const double v[13] = {qz, rx, ry, rz, rxx, ryy, rzz, rxx * 2.0, ryy * 2.0, rzz * 2.0,
                          rxy * 2.0, ryz * 2.0, rzx * 2.0};
    for (int k = 0; k < 13; ++k) P[k] = v[k];
}

// Table checks of neartb (DOUB): partners per atom and class in the EL-only
// table (itbEL), the EL+vdW table (itb) and the second-nearest LJ table
// against their capacities (ier 51, neartb.f90:368-400), and the excluded
// pairs the tables hold (ier 99 unless all, neartb.f90:402-453).
__global__ void k_cmm_check(CmmArgs A) {
#pragma clang fp contract(off)
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= A.n || (A.err[0] & 1)) return;
    const SysP &P = A.sp[0];
    const int a = A.sorig[p], tf = A.sflag[p];
    const int fi = (tf >> 13) & 1, vi = (tf >> 12) & 1;
    const int nf = A.nf, nv = A.nv1;
    const int cell = A.key[a] / (2 * nf);
    const int cx = cell % nv, cy = (cell / nv) % nv, cz = cell / (nv * nv);
    int nel[3] = {0, 0, 0}, nvd[3] = {0, 0, 0}, nvw[3] = {0, 0, 0};
    for (int z = max(cz - 1, 0); z <= min(cz + 1, nv - 1); ++z)
        for (int y = max(cy - 1, 0); y <= min(cy + 1, nv - 1); ++y)
            for (int x = max(cx - 1, 0); x <= min(cx + 1, nv - 1); ++x) {
                const int c2 = cmm_lin(nv, x, y, z);
                for (int fj = 0; fj < nf; ++fj) {
                    const int k = cmm_cls(P, fi, fj);
                    const int q = (c2 * nf + fj) * 2;
                    const int own = (c2 == cell && fj == fi) ? 1 : 0;
                    const int n0 = A.kcnt[q], n1 = A.kcnt[q + 1];
                    if (!vi) {
                        nel[k] += n0 + n1 - own;
                    } else {
                        nel[k] += n0;
                        nvd[k] += n1 - own;
                    }
                }
            }
    const double xi = A.sx0[3 * p], yi = A.sx0[3 * p + 1], zi = A.sx0[3 * p + 2];
    if (vi && A.idelay > 1) {
        const int d = A.idelay;
        for (int z = max(cz - d, 0); z <= min(cz + d, nv - 1); ++z)
            for (int y = max(cy - d, 0); y <= min(cy + d, nv - 1); ++y)
                for (int x = max(cx - d, 0); x <= min(cx + d, nv - 1); ++x) {
                    if (abs(z - cz) <= 1 && abs(y - cy) <= 1 && abs(x - cx) <= 1) continue;
                    const int c2 = cmm_lin(nv, x, y, z);
                    for (int fj = 0; fj < nf; ++fj) {
                        const int k = cmm_cls(P, fi, fj);
                        const int q = (c2 * nf + fj) * 2 + 1;
                        for (int j = A.kstart[q]; j < A.kstart[q] + A.kcnt[q]; ++j) {
                            const double d0 = xi - A.sx0[3 * j], d1 = yi - A.sx0[3 * j + 1],
                                         d2 = zi - A.sx0[3 * j + 2];
                            if (d0 * d0 + d1 * d1 + d2 * d2 < A.rlim2) nvw[k]++;
                        }
                    }
                }
    }
    bool over = false;
    for (int k = 0; k < 3; ++k) over = over || nel[k] > A.capel || nvd[k] > A.capvd || nvw[k] > A.capvw;
    if (over) atomicOr(A.err, 2);
    int found = 0;
    for (int e = A.xptr[a]; e < A.xptr[a + 1]; ++e) {
        const int b = A.xidx[e];
        const int c2 = A.key[b] / (2 * nf);
        const int dx = abs(c2 % nv - cx), dy = abs((c2 / nv) % nv - cy), dz = abs(c2 / (nv * nv) - cz);
        const int cheb = max(dx, max(dy, dz));
        if (cheb <= 1) {
            ++found;
        } else if (A.idelay > 1 && vi && ((A.tflag[b] >> 12) & 1) && cheb <= A.idelay) {
            const int pb = A.spos[b];
            const double d0 = xi - A.sx0[3 * pb], d1 = yi - A.sx0[3 * pb + 1], d2 = zi - A.sx0[3 * pb + 2];
            if (d0 * d0 + d1 * d1 + d2 * d2 < A.rlim2) ++found;
        }
    }
    if (found) atomicAdd(A.err + 1, found);
}

// ------------------------------------------------------------------ per step

__global__ void k_cmm_gather(CmmArgs A) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= A.n) return;
    const int a = A.sorig[p];
    A.sx[p] = make_double4(A.x[3 * a], A.x[3 * a + 1], A.x[3 * a + 2], A.chg[a]);
}

__device__ __forceinline__ bool cmm_excluded(const CmmArgs &A, int ai, int aj) {
    for (int e = A.xptr[ai]; e < A.xptr[ai + 1]; ++e)
        if (A.xidx[e] == aj) return true;
    return false;
}

// Lanes = (replica r, i atom li) with R replicas of nic atoms in one wave.
struct CmmLanes {
    int R, nic, r, li;
    __device__ CmmLanes(int ni, int lane) {
        nic = min(ni, 64);
        R = 1;
        while (2 * R * nic <= 64) R *= 2;
        r = lane / nic;
        li = lane - r * nic;
    }
    __device__ bool active() const { return r < R; }
};

// Per-lane sums of one i atom: gradient per class and energies per class.
template <bool MULTI>
struct CmmAcc {
    double g[MULTI ? 3 : 1][3];
    double ee[MULTI ? 3 : 1], vv[MULTI ? 3 : 1];
    __device__ void zero() {
        for (int k = 0; k < (MULTI ? 3 : 1); ++k) {
            g[k][0] = g[k][1] = g[k][2] = 0.0;
            ee[k] = vv[k] = 0.0;
        }
    }
    // md_run's sums: e, then gx + dx * cc (enear_CMM_doub)
    __device__ void add(int k, double dx, double dy, double dz, double cc, double e, double v) {
#pragma clang fp contract(off)
        if constexpr (MULTI) {
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                const bool on = m == k;
                g[m][0] = g[m][0] + (on ? dx * cc : 0.0);
                g[m][1] = g[m][1] + (on ? dy * cc : 0.0);
                g[m][2] = g[m][2] + (on ? dz * cc : 0.0);
                ee[m] = ee[m] + (on ? e : 0.0);
                vv[m] = vv[m] + (on ? v : 0.0);
            }
        } else {
            g[0][0] = g[0][0] + dx * cc;
            g[0][1] = g[0][1] + dy * cc;
            g[0][2] = g[0][2] + dz * cc;
            ee[0] = ee[0] + e;
            vv[0] = vv[0] + v;
        }
    }
};

// Sum the replicas of every i atom in replica order through LDS and store
// (or add, for the second-nearest pass) the result of the sorted atom p.
template <bool MULTI, bool E>
__device__ void cmm_reduce_store(const CmmArgs &A, const CmmAcc<MULTI> &acc, const CmmLanes &L, int p, bool add,
                                 double (*s)[15]) {
#pragma clang fp contract(off)
    constexpr int NC = MULTI ? 3 : 1;
    const int lane = threadIdx.x;
    for (int k = 0; k < NC; ++k) {
        s[lane][3 * k] = acc.g[k][0];
        s[lane][3 * k + 1] = acc.g[k][1];
        s[lane][3 * k + 2] = acc.g[k][2];
        s[lane][9 + k] = acc.ee[k];
        s[lane][12 + k] = acc.vv[k];
    }
    __syncthreads();
    if (L.r == 0 && L.li < L.nic && p >= 0) {
        double t[15];
        for (int m = 0; m < 15; ++m) t[m] = 0.0;
        for (int r = 0; r < L.R; ++r) {
            const int ln = r * L.nic + L.li;
            for (int k = 0; k < NC; ++k) {
                for (int d = 0; d < 3; ++d) t[3 * k + d] = t[3 * k + d] + s[ln][3 * k + d];
                t[9 + k] = t[9 + k] + s[ln][9 + k];
                t[12 + k] = t[12 + k] + s[ln][12 + k];
            }
        }
        double *G = A.gnb + (size_t)p * 9;
        double *En = A.enb + (size_t)p * 6;
        for (int k = 0; k < 3; ++k) {
            for (int d = 0; d < 3; ++d) {
                const double v = k < NC ? t[3 * k + d] : 0.0;
                G[3 * k + d] = add ? G[3 * k + d] + v : v;
            }
            if (E) {
                const double e = k < NC ? t[9 + k] : 0.0, v = k < NC ? t[12 + k] : 0.0;
                En[k] = add ? En[k] + e : e;
                En[3 + k] = add ? En[3 + k] + v : v;
            }
        }
    }
    __syncthreads();
}

// Near field: every pair of the cell and its 26 neighbours (clipped at the
// field's faces).  One wavefront per level-1 cell.
template <bool MULTI, bool E>
__global__ void __launch_bounds__(64) k_cmm_near(CmmArgs A) {
#pragma clang fp contract(off)
    __shared__ double s[64][15];
    const int cell = blockIdx.x, nf = A.nf, nv = A.nv1;
    const int cs = A.kstart[cell * nf * 2], ce = A.kstart[(cell + 1) * nf * 2];
    if (ce == cs) return;
    const SysP &P = A.sp[0];
    const int cx = cell % nv, cy = (cell / nv) % nv, cz = cell / (nv * nv);
    const int nt = A.ntyp;
    for (int i0 = cs; i0 < ce; i0 += 64) {
        const CmmLanes L(ce - i0, threadIdx.x);
        const int p = i0 + L.li;
        const bool act = L.active();
        const double4 xi = act ? A.sx[p] : make_double4(0.0, 0.0, 0.0, 0.0);
        const int fi = act ? A.sflag[p] : 0;
        const int ti = fi & 0xfff, vi = (fi >> 12) & 1, ci = (fi >> 13) & 1;
        const int moli = act ? A.smol[p] : -1;
        const int ai = act ? A.sorig[p] : 0;
        CmmAcc<MULTI> acc;
        acc.zero();
        for (int z = max(cz - 1, 0); z <= min(cz + 1, nv - 1); ++z)
            for (int y = max(cy - 1, 0); y <= min(cy + 1, nv - 1); ++y)
                for (int x = max(cx - 1, 0); x <= min(cx + 1, nv - 1); ++x) {
                    const int c2 = cmm_lin(nv, x, y, z);
                    const int js = A.kstart[c2 * nf * 2], je = A.kstart[(c2 + 1) * nf * 2];
                    if (!act) continue;
                    for (int j = js + L.r; j < je; j += L.R) {
                        if (j == p) continue;
                        if (A.smol[j] == moli && cmm_excluded(A, ai, A.sorig[j])) continue;
                        const double4 xj = A.sx[j];
                        const int fj = A.sflag[j];
                        const double dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
                        const double d2i = 1.0 / (dx * dx + dy * dy + dz * dz);
                        const double e = xi.w * xj.w * sqrt(d2i);
                        double cc = e, v = 0.0;
                        if (vi & (fj >> 12)) {
                            const double2 c = A.lj[(fj & 0xfff) * nt + ti];
                            const double d6i = d2i * d2i * d2i;
                            const double v1 = c.x * d6i, v2 = c.y * d6i * d6i;
                            v = v2 - v1;
                            cc = e - 6.0 * v1 + 12.0 * v2;
                        }
                        cc = cc * d2i;
                        acc.add(MULTI ? cmm_cls(P, ci, (fj >> 13) & 1) : 0, dx, dy, dz, cc, e, v);
                    }
                }
        cmm_reduce_store<MULTI, E>(A, acc, L, act ? p : -1, false, s);
    }
}

// LJ of the second-nearest cells (idelay > 1): vdW pairs within rlim at the
// rebuild, in cells at Chebyshev distance 2 .. idelay (neartb.f90:321-366,
// enear's third table).  One wavefront per cell over its vdW atoms.
template <bool MULTI, bool E>
__global__ void __launch_bounds__(64) k_cmm_lj2(CmmArgs A) {
#pragma clang fp contract(off)
    __shared__ double s[64][15];
    __shared__ int vlist[256];
    const int cell = blockIdx.x, nf = A.nf, nv = A.nv1;
    int nvi = 0;
    for (int f = 0; f < nf; ++f) nvi += A.kcnt[(cell * nf + f) * 2 + 1];
    if (nvi == 0) return;
    const SysP &P = A.sp[0];
    const int cx = cell % nv, cy = (cell / nv) % nv, cz = cell / (nv * nv);
    const int nt = A.ntyp, d = A.idelay;
    for (int i0 = 0; i0 < nvi; i0 += 256) {
        // this block's vdW atoms (sorted positions), fragment by fragment
        if (threadIdx.x == 0) {
            int w = 0, skip = i0;
            for (int f = 0; f < nf && w < 256; ++f) {
                const int q = (cell * nf + f) * 2 + 1;
                for (int t = 0; t < A.kcnt[q] && w < 256; ++t) {
                    if (skip > 0) { --skip; continue; }
                    vlist[w++] = A.kstart[q] + t;
                }
            }
        }
        __syncthreads();
        const int nchunk = min(256, nvi - i0);
        for (int c0 = 0; c0 < nchunk; c0 += 64) {
            const CmmLanes L(nchunk - c0, threadIdx.x);
            const bool act = L.active();
            const int p = act ? vlist[c0 + L.li] : -1;
            const double4 xi = act ? A.sx[p] : make_double4(0.0, 0.0, 0.0, 0.0);
            const double x0 = act ? A.sx0[3 * p] : 0.0, y0 = act ? A.sx0[3 * p + 1] : 0.0,
                         z0 = act ? A.sx0[3 * p + 2] : 0.0;
            const int fi = act ? A.sflag[p] : 0;
            const int ti = fi & 0xfff, ci = (fi >> 13) & 1;
            const int moli = act ? A.smol[p] : -1;
            const int ai = act ? A.sorig[p] : 0;
            CmmAcc<MULTI> acc;
            acc.zero();
            if (act) {
                for (int z = max(cz - d, 0); z <= min(cz + d, nv - 1); ++z)
                    for (int y = max(cy - d, 0); y <= min(cy + d, nv - 1); ++y)
                        for (int x = max(cx - d, 0); x <= min(cx + d, nv - 1); ++x) {
                            if (abs(z - cz) <= 1 && abs(y - cy) <= 1 && abs(x - cx) <= 1) continue;
                            const int c2 = cmm_lin(nv, x, y, z);
                            for (int fj = 0; fj < nf; ++fj) {
                                const int q = (c2 * nf + fj) * 2 + 1;
                                const int js = A.kstart[q], je = js + A.kcnt[q];
                                for (int j = js + L.r; j < je; j += L.R) {
                                    const double e0 = x0 - A.sx0[3 * j], e1 = y0 - A.sx0[3 * j + 1],
                                                 e2 = z0 - A.sx0[3 * j + 2];
                                    if (!(e0 * e0 + e1 * e1 + e2 * e2 < A.rlim2)) continue;
                                    if (A.smol[j] == moli && cmm_excluded(A, ai, A.sorig[j])) continue;
                                    const double4 xj = A.sx[j];
                                    const int fj2 = A.sflag[j];
                                    const double dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
                                    const double2 c = A.lj[(fj2 & 0xfff) * nt + ti];
                                    const double d2i = 1.0 / (dx * dx + dy * dy + dz * dz);
                                    const double d6i = d2i * d2i * d2i;
                                    const double v1 = c.x * d6i, v2 = c.y * d6i * d6i;
                                    const double cc = (-6.0 * v1 + 12.0 * v2) * d2i;
                                    acc.add(MULTI ? cmm_cls(P, ci, (fj2 >> 13) & 1) : 0, dx, dy, dz, cc, 0.0,
                                            v2 - v1);
                                }
                            }
                        }
            }
            cmm_reduce_store<MULTI, E>(A, acc, L, p, true, s);
        }
        __syncthreads();
    }
}

// Far field (compute_particle_multipole, cmmdo.f90:164-288): for every atom,
// per source fragment and level, the cells of that level in the far list of
// its ancestor (mk_cm: the children of the ancestor's parent's neighbours
// less the ancestor's own neighbours; all cells less the neighbours at the
// top level) that hold atoms of the fragment.  The per-atom sums run in
// md_run's order; the gradient is scaled by chgmod2 = ccoef q at the end.
template <bool MULTI>
__global__ void __launch_bounds__(64) k_cmm_far(CmmArgs A) {
#pragma clang fp contract(off)
    const int cell = blockIdx.x, nf = A.nf, nv1 = A.nv1;
    const int cs = A.kstart[cell * nf * 2], ce = A.kstart[(cell + 1) * nf * 2];
    if (ce == cs) return;
    const SysP &P = A.sp[0];
    constexpr int NC = MULTI ? 3 : 1;
    for (int i0 = cs; i0 < ce; i0 += 64) {
        const int p = i0 + (int)threadIdx.x;
        const bool act = p < ce;
        const int a = act ? A.sorig[p] : 0;
        const double4 xi = act ? A.sx[p] : make_double4(0.0, 0.0, 0.0, 0.0);
        const int fi = act ? A.sflag[p] : 0;
        const int k = (fi >> 13) & 1;
        double G[NC][3], Ec[NC];
        for (int m = 0; m < NC; ++m) { G[m][0] = G[m][1] = G[m][2] = 0.0; Ec[m] = 0.0; }
        const double qa = act ? A.q[a] : 0.0;
        for (int kk = 0; kk < nf; ++kk) {
            const int ncls = MULTI ? cmm_cls(P, kk, k) : 0;
            int ix = cell % nv1, iy = (cell / nv1) % nv1, iz = cell / (nv1 * nv1);
            for (int l = 0; l < A.nlev; ++l) {
                if (l > 0) { ix >>= 1; iy >>= 1; iz >>= 1; }
                const int mm = A.nv[l];
                int lo[3], hi[3];
                const int ii[3] = {ix, iy, iz};
                for (int t = 0; t < 3; ++t) {
                    if (l == A.nlev - 1) {
                        lo[t] = 0; hi[t] = mm - 1;
                    } else {
                        const int m1 = A.nv[l + 1], pj = ii[t] >> 1;
                        lo[t] = max(pj - 1, 0) * 2;
                        hi[t] = min(pj + 1, m1 - 1) * 2 + 1;
                    }
                }
                const double *cgl = A.cg + A.cgoff[l];
                for (int jz = lo[2]; jz <= hi[2]; ++jz)
                    for (int jy = lo[1]; jy <= hi[1]; ++jy)
                        for (int jx = lo[0]; jx <= hi[0]; ++jx) {
                            if (abs(jz - iz) <= 1 && abs(jy - iy) <= 1 && abs(jx - ix) <= 1) continue;
                            const int jc = cmm_lin(mm, jx, jy, jz);
                            if (A.npc[A.pcoff[l] + jc * nf + kk] == 0) continue;
                            const double *q = A.pot + (size_t)(A.pcoff[l] + jc * nf + kk) * 13;
                            const double dx = xi.x - (cgl[jx] + A.c0[0]), dy = xi.y - (cgl[jy] + A.c0[1]),
                                         dz = xi.z - (cgl[jz] + A.c0[2]);
                            const double dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
                            const double d2i = 1.0 / (dx2 + dy2 + dz2);
                            const double di = sqrt(d2i), d3i = di * d2i, d5i = d3i * d2i;
                            const double dxy = dx * dy, dyz = dy * dz, dzx = dz * dx;
                            const double aa1 = q[1] * dx + q[2] * dy + q[3] * dz;
                            const double aa2 = q[4] * dx2 + q[5] * dy2 + q[6] * dz2 + q[10] * dxy + q[11] * dyz +
                                               q[12] * dzx;
                            const double e = q[0] * di + aa1 * d3i + aa2 * d5i;
                            const double bbx = (q[0] * dx - q[1]) * d3i, bby = (q[0] * dy - q[2]) * d3i,
                                         bbz = (q[0] * dz - q[3]) * d3i;
                            const double cc2x = q[7] * dx + q[10] * dy + q[12] * dz;
                            const double cc2y = q[10] * dx + q[8] * dy + q[11] * dz;
                            const double cc2z = q[12] * dx + q[11] * dy + q[9] * dz;
                            const double cc4 = aa1 * 3.0 + aa2 * d2i * 5.0;
                            const double gx = -bbx + (cc2x - cc4 * dx) * d5i;
                            const double gy = -bby + (cc2y - cc4 * dy) * d5i;
                            const double gz = -bbz + (cc2z - cc4 * dz) * d5i;
                            const double ea = qa * A.ccoef * e;
#pragma unroll
                            for (int m = 0; m < NC; ++m) {
                                const bool on = m == ncls;
                                G[m][0] = G[m][0] + (on ? gx : 0.0);
                                G[m][1] = G[m][1] + (on ? gy : 0.0);
                                G[m][2] = G[m][2] + (on ? gz : 0.0);
                                Ec[m] = Ec[m] + (on ? ea : 0.0);
                            }
                        }
            }
        }
        if (act) {
            const double q2 = A.ccoef * qa;
            for (int m = 0; m < 3; ++m) {
                for (int t = 0; t < 3; ++t) A.gwk[((size_t)a * 3 + m) * 3 + t] = m < NC ? G[m][t] * q2 : 0.0;
                A.efar[(size_t)a * 3 + m] = m < NC ? Ec[m] : 0.0;
            }
        }
    }
}

// grad = sum over classes of factor * ((0 - near) + far) into the slot
// gradient (slot = atom); the energies of the step per atom.
template <bool E>
__global__ void k_cmm_combine(CmmArgs A) {
#pragma clang fp contract(off)
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= A.n) return;
    const int p = A.spos[a];
    const double *G = A.gnb + (size_t)p * 9, *W = A.gwk + (size_t)a * 9;
    double g[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < A.nfrg; ++k) {
        const double f = A.efac[k];
        for (int d = 0; d < 3; ++d) g[d] = g[d] + f * ((0.0 - G[3 * k + d]) + W[3 * k + d]);
    }
    for (int d = 0; d < 3; ++d) A.gslot[3 * (size_t)a + d] = g[d];
    if (E) {
        const double *En = A.enb + (size_t)p * 6;
        for (int k = 0; k < 3; ++k) {
            A.eatom[6 * (size_t)a + k] = En[k] + A.efar[3 * (size_t)a + k];
            A.eatom[6 * (size_t)a + 3 + k] = En[3 + k];
        }
    }
}
