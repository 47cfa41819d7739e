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
//   npc, pot: atoms of every (cell, fragment) at every level (celoth) and
//   its record of 16 doubles: the shifted cell centre and the 13 multipole
//   coefficients (celpol).
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
    int pcoff[JP_CMM_MAXLEV];   // offset of level l's (cell, fragment) entries in npc (x 16 in pot)
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
    double *gwkp;               // [part][a][12] k_cmm_far's partial sums (x, y, z, energy per class)
    double *efar;               // [a][3] far-field energy per class (kept likewise)
    double *gslot;              // [a][3] combined gradient
    double *eatom;              // [a][6] elec (near + far), vdW per class; summed and halved by the host
    unsigned long long *xmask;  // [a] excluded partners b with b - a in [-32, 31] (bit b - a + 32; bit 32: a)
    int *xwide;                 // [a] 1: some excluded partner beyond that window
    int *l2cnt, *l2ptr, *l2idx; // second-nearest LJ table per sorted atom (fixed between rebuilds)
    int *vflag, *vpos, *vlist;  // the sorted atoms with such a table; how many at vpos[n]
    // the vdW atoms alone, in the sorted order (second-nearest candidates):
    // kvstart[key] how many precede the key, then per vdW atom its
    // coordinates at the rebuild, flags, atom number, molecule, sorted position
    int *kvcnt, *kvstart;
    double4 *vx0;
    int *vf, *va, *vm, *vp;
};

__device__ __forceinline__ int cmm_lin(int nv, int x, int y, int z) { return (z * nv + y) * nv + x; }

// Energy class of a pair of cluster classes (n_matrix - 1).
__device__ __forceinline__ int cmm_cls(const SysP &P, int ci, int cj) { return P.cls[ci * 2 + cj]; }

// 1/sqrt(x) for the pair kernels (x a squared distance, far from 0 and
// infinity): the hardware estimate and two Newton steps, each doubling its
// correct bits, to within a few units in the last place of FP64.  md_run
// forms 1/r^2 and sqrt of it; the pair terms here are within 1e-15 of those.
__device__ __forceinline__ double cmm_rsqrt(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double hx = 0.5 * x;
    y = y * (1.5 - hx * y * y);
    y = y * (1.5 - hx * y * y);
    return y;
}

// ------------------------------------------------------------------ scans

// Exclusive scan of n ints in three launches: k_scan_tile scans tiles of
// 1024 (4 per thread) and stores their totals, k_scan (one block) scans the
// totals in place, k_scan_add adds each tile's offset; out[n] receives the
// total.  tsum holds ceil(n / 1024) + 1 ints.
__global__ void __launch_bounds__(256) k_scan_tile(int n, const int *__restrict__ in, int *__restrict__ out,
                                                   int *__restrict__ tsum) {
    __shared__ int wsum[4];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int i0 = blockIdx.x * 1024 + 4 * t;
    int v[4], s = 0;
    for (int k = 0; k < 4; ++k) {
        v[k] = i0 + k < n ? in[i0 + k] : 0;
        s += v[k];
    }
    int x = s;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int base = 0;
    for (int k = 0; k < w; ++k) base += wsum[k];
    int e = base + x - s;
    for (int k = 0; k < 4; ++k) {
        if (i0 + k < n) out[i0 + k] = e;
        e += v[k];
    }
    if (t == 255) tsum[blockIdx.x] = base + x;
}

__global__ void k_scan_add(int n, int *__restrict__ out, const int *__restrict__ tsum, int ntile) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += tsum[i >> 10];
    if (i == 0) out[n] = tsum[ntile];
}

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

// The vdW atoms alone (second-nearest candidates): their count per key, and,
// after its scan, their copy in the sorted order.
__global__ void k_cmm_kv(CmmArgs A) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < A.nkey) A.kvcnt[k] = (k & 1) ? A.kcnt[k] : 0;
}

__global__ void k_cmm_vcopy(CmmArgs A) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= A.n || (A.err[0] & 1)) return;
    const int a = A.sorig[p], k = A.key[a];
    if (!(k & 1)) return;
    const int q = A.kvstart[k] + (p - A.kstart[k]);
    A.vx0[q] = make_double4(A.sx0[3 * p], A.sx0[3 * p + 1], A.sx0[3 * p + 2], 0.0);
    A.vf[q] = A.sflag[p];
    A.va[q] = a;
    A.vm[q] = A.smol[p];
    A.vp[q] = p;
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
    double *P = A.pot + (size_t)(A.pcoff[l] + q) * 16;
    const double *cgl = A.cg + A.cgoff[l];
    const double gx = cgl[ix] + A.c0[0], gy = cgl[iy] + A.c0[1], gz = cgl[iz] + A.c0[2];
    P[0] = gx; P[1] = gy; P[2] = gz;
    P += 3;
    if (A.npc[A.pcoff[l] + q] == 0) {
        for (int k = 0; k < 13; ++k) P[k] = 0.0;
        return;
    }
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
        const double *C = A.pot + (size_t)qc * 16 + 3;
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

// Excluded pair (or i = j): the window mask of atom i (partners b with
// b - ai in [-32, 31], bit 32 being i itself), else, for atoms with
// partners beyond the window, the partner list when both lie in one molecule.
__device__ __forceinline__ bool cmm_excl(const CmmArgs &A, int ai, int aj, unsigned long long mask, bool wide,
                                         int moli, int molj) {
    const unsigned d = (unsigned)(aj - ai + 32);
    if (d < 64u && ((mask >> d) & 1ull)) return true;
    return wide && molj == moli && cmm_excluded(A, ai, aj);
}

// Set-up: the exclusion window masks.
__global__ void k_cmm_xmask(CmmArgs A) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= A.n) return;
    unsigned long long m = 1ull << 32;
    int wide = 0;
    for (int e = A.xptr[a]; e < A.xptr[a + 1]; ++e) {
        const unsigned d = (unsigned)(A.xidx[e] - a + 32);
        if (d < 64u) m |= 1ull << d;
        else wide = 1;
    }
    A.xmask[a] = m;
    A.xwide[a] = wide;
}

// Table checks of neartb (DOUB): partners per atom and class in the EL-only
// table (itbEL) and the EL+vdW table (itb) against their capacities (ier 51,
// neartb.f90:368-400), and the excluded pairs the tables hold (ier 99 unless
// all, neartb.f90:402-453).  The second-nearest LJ table is k_cmm_l2's.
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
    int nel[3] = {0, 0, 0}, nvd[3] = {0, 0, 0};
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
    A.l2cnt[p] = 0;   // k_cmm_l2 sets those of the vdW atoms
    bool over = false;
    for (int k = 0; k < 3; ++k) over = over || nel[k] > A.capel || nvd[k] > A.capvd;
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

// In-place exclusive scan of a[0, n) in LDS by a block of 256 threads
// (n <= 2048, 8 entries per thread); a[n] receives the total.  Every thread
// of the block calls it; it ends with a barrier.
__device__ void cmm_block_scan(int *a, int n) {
    __shared__ int ws[4];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, b = 8 * t;
    int v[8], s = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        v[k] = b + k < n ? a[b + k] : 0;
        s += v[k];
    }
    int x = s;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) ws[w] = x;
    __syncthreads();
    int base = 0;
    for (int k = 0; k < w; ++k) base += ws[k];
    int e = base + x - s;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        if (b + k < n) a[b + k] = e;
        e += v[k];
    }
    if (t == 255) a[n] = base + x;
    __syncthreads();
}

// Per-thread sums of one i atom: gradient per class and energies per class.
template <bool MULTI>
struct CmmAcc {
    static constexpr int NC = MULTI ? 3 : 1;
    static constexpr int NV = 5 * NC;   // g[k][3], ee[k], vv[k]
    double v[NV];
    __device__ void zero() {
#pragma unroll
        for (int m = 0; m < NV; ++m) v[m] = 0.0;
    }
    // md_run's sums: e, then gx + dx * cc (enear_CMM_doub)
    __device__ void add(int k, double dx, double dy, double dz, double cc, double e, double vw) {
#pragma unroll
        for (int m = 0; m < NC; ++m) {
            const bool on = !MULTI || m == k;
            v[3 * m] = v[3 * m] + (on ? dx * cc : 0.0);
            v[3 * m + 1] = v[3 * m + 1] + (on ? dy * cc : 0.0);
            v[3 * m + 2] = v[3 * m + 2] + (on ? dz * cc : 0.0);
            v[3 * NC + m] = v[3 * NC + m] + (on ? e : 0.0);
            v[4 * NC + m] = v[4 * NC + m] + (on ? vw : 0.0);
        }
    }
};

#define CMM_NT 256     // threads per block of k_cmm_near and k_cmm_far
#define CMM_JCH 512    // j atoms staged in LDS per pass

// Near field: every pair of the cell and its 26 neighbours (clipped at the
// field's faces).  One block per level-1 cell: the j atoms of the 27 cells
// (nine runs of x-adjacent cells, contiguous in the sorted order) are staged
// through LDS in chunks; thread t takes i atom t % nic and, as replica
// t / nic of R = 256 / nic, every R-th staged j; the replicas are summed in
// replica order.
template <bool MULTI, bool E>
__global__ void __launch_bounds__(CMM_NT) k_cmm_near(CmmArgs A) {
    using Acc = CmmAcc<MULTI>;
    constexpr int NV = Acc::NV, NC = Acc::NC;
    constexpr int SJ = 4 * CMM_JCH + 3 * CMM_JCH / 2, SR = CMM_NT * NV;
    __shared__ double s_buf[SJ > SR ? SJ : SR];
    __shared__ int s_run[9], s_off[10];
    extern __shared__ double2 s_lj[];
    double *jx = s_buf, *jy = jx + CMM_JCH, *jz = jy + CMM_JCH, *jq = jz + CMM_JCH;
    int2 *jfa = (int2 *)(jq + CMM_JCH);   // (flags, atom number)
    int *jm = (int *)(jfa + CMM_JCH);
    const int cell = blockIdx.x, nf = A.nf, nv = A.nv1, t = threadIdx.x;
    const int cs = A.kstart[cell * nf * 2], ce = A.kstart[(cell + 1) * nf * 2];
    if (ce == cs) return;
    const SysP &P = A.sp[0];
    const int nt = A.ntyp;
    for (int k = t; k < nt * nt; k += CMM_NT) s_lj[k] = A.lj[k];
    const int cx = cell % nv, cy = (cell / nv) % nv, cz = cell / (nv * nv);
    if (t < 9) {   // the nine runs (z, y) of x-adjacent cells, one per thread; empty off the field
        const int z = cz + t / 3 - 1, y = cy + t % 3 - 1;
        int b = 0, e = 0;
        if (z >= 0 && z < nv && y >= 0 && y < nv) {
            b = A.kstart[cmm_lin(nv, max(cx - 1, 0), y, z) * nf * 2];
            e = A.kstart[(cmm_lin(nv, min(cx + 1, nv - 1), y, z) + 1) * nf * 2];
        }
        s_run[t] = b;
        s_off[t] = e - b;
    }
    __syncthreads();
    if (t == 0) {
        int off = 0;
        for (int k = 0; k < 9; ++k) {
            const int len = s_off[k];
            s_off[k] = off;
            off += len;
        }
        s_off[9] = off;
    }
    __syncthreads();
    const int J = s_off[9];
    for (int i0 = cs; i0 < ce; i0 += CMM_NT) {
        const int nic = min(CMM_NT, ce - i0);
        const int R = CMM_NT / nic, r = t / nic, li = t - r * nic;
        const bool act = r < R;
        const int p = i0 + li;
        const double4 xi = act ? A.sx[p] : make_double4(0.0, 0.0, 0.0, 0.0);
        const int fi = act ? A.sflag[p] : 0;
        const int ti = fi & 0xfff, vi = (fi >> 12) & 1, ci = (fi >> 13) & 1;
        const int ai = act ? A.sorig[p] : 0;
        const int moli = act ? A.smol[p] : -1;
        const unsigned long long mask = act ? A.xmask[ai] : 0ull;
        const bool wide = act && A.xwide[ai];
        Acc acc;
        acc.zero();
        for (int q0 = 0; q0 < J; q0 += CMM_JCH) {
            const int nq = min(CMM_JCH, J - q0);
            __syncthreads();
            for (int s = t; s < nq; s += CMM_NT) {
                const int g = q0 + s;
                int k = 0;
                while (g >= s_off[k + 1]) ++k;
                const int pj = s_run[k] + (g - s_off[k]);
                const double4 v = A.sx[pj];
                jx[s] = v.x; jy[s] = v.y; jz[s] = v.z; jq[s] = v.w;
                jfa[s] = make_int2(A.sflag[pj], A.sorig[pj]);
                jm[s] = A.smol[pj];
            }
            __syncthreads();
            if (act) {
                for (int s = r; s < nq; s += R) {
                    const int2 fa = jfa[s];
                    // excluded pairs and i itself: the window mask, else the list
                    const unsigned dw = (unsigned)(fa.y - ai + 32);
                    if (dw < 64u && ((mask >> dw) & 1ull)) continue;
                    if (wide && jm[s] == moli && cmm_excluded(A, ai, fa.y)) continue;
                    const int fj = fa.x;
                    const double dx = xi.x - jx[s], dy = xi.y - jy[s], dz = xi.z - jz[s];
                    const double ri = cmm_rsqrt(dx * dx + dy * dy + dz * dz);
                    const double d2i = ri * ri;
                    const double e = xi.w * jq[s] * ri;
                    double cc = e, vw = 0.0;
                    if (vi & (fj >> 12)) {
                        const double2 c = s_lj[(fj & 0xfff) * nt + ti];
                        const double d6i = d2i * d2i * d2i;
                        const double v1 = c.x * d6i, v2 = c.y * d6i * d6i;
                        vw = v2 - v1;
                        cc = e - 6.0 * v1 + 12.0 * v2;
                    }
                    cc = cc * d2i;
                    acc.add(MULTI ? cmm_cls(P, ci, (fj >> 13) & 1) : 0, dx, dy, dz, cc, e, vw);
                }
            }
        }
        __syncthreads();
        double *red = s_buf;
#pragma unroll
        for (int m = 0; m < NV; ++m) red[m * CMM_NT + t] = acc.v[m];
        __syncthreads();
        if (act && r == 0) {
            double tot[NV];
#pragma unroll
            for (int m = 0; m < NV; ++m) tot[m] = 0.0;
            for (int rr = 0; rr < R; ++rr)
#pragma unroll
                for (int m = 0; m < NV; ++m) tot[m] = tot[m] + red[m * CMM_NT + rr * nic + li];
            double *G = A.gnb + (size_t)p * 9;
            for (int k = 0; k < 3; ++k)
                for (int d = 0; d < 3; ++d) G[3 * k + d] = k < NC ? tot[3 * k + d] : 0.0;
            if (E) {
                double *En = A.enb + (size_t)p * 6;
                for (int k = 0; k < 3; ++k) {
                    En[k] = k < NC ? tot[3 * NC + k] : 0.0;
                    En[3 + k] = k < NC ? tot[4 * NC + k] : 0.0;
                }
            }
        }
    }
}

#define CMM_L2CH 512    // candidates staged in LDS per pass of k_cmm_l2
#define CMM_MAXRUN 2048 // run slots of a shell, 2 (2 idelay + 1)^2 (idelay <= 15)

// The cells at Chebyshev distance 2 .. d of cell (cx, cy, cz) as runs of
// x-adjacent cells (contiguous ranges of the vdW atoms), two slots per row
// (z, y), built by the whole block: s_run[k] the first vdW atom of slot k,
// s_off[k] its offset in the concatenation (empty slots have none); returns
// the slot count, s_off[count] the total.
__device__ int cmm_shell_runs(const CmmArgs &A, int cx, int cy, int cz, int d, int *s_run, int *s_off) {
    const int nv = A.nv1, nf = A.nf, w = 2 * d + 1, nrow = w * w;
    const int x0 = max(cx - d, 0), x1 = min(cx + d, nv - 1);
    for (int k = threadIdx.x; k < nrow; k += blockDim.x) {
        const int z = cz - d + k / w, y = cy - d + k % w;
        int xa[2] = {0, 0}, xb[2] = {-1, -1};
        if (z >= 0 && z < nv && y >= 0 && y < nv) {
            if (abs(z - cz) <= 1 && abs(y - cy) <= 1) {
                xa[0] = x0; xb[0] = cx - 2;
                xa[1] = cx + 2; xb[1] = x1;
            } else {
                xa[0] = x0; xb[0] = x1;
            }
        }
        for (int h = 0; h < 2; ++h) {
            int b = 0, len = 0;
            if (xa[h] <= xb[h]) {
                b = A.kvstart[cmm_lin(nv, xa[h], y, z) * nf * 2];
                len = A.kvstart[(cmm_lin(nv, xb[h], y, z) + 1) * nf * 2] - b;
            }
            s_run[2 * k + h] = b;
            s_off[2 * k + h] = len;
        }
    }
    __syncthreads();
    cmm_block_scan(s_off, 2 * nrow);
    return 2 * nrow;
}

// The second-nearest LJ tables (neartb.f90:321-366): for every vdW atom the
// vdW atoms of the cells at Chebyshev distance 2 .. idelay within rlim at the
// rebuild (sx0), less its excluded pairs.  One block per cell: the
// candidates (the vdW atoms of the shell's cells, from the vdW copy) staged
// through LDS, thread t takes vdW atom t % nic and, as
// replica r = t / nic of R = 256 / nic, every R-th candidate.  COUNT: the
// partners per class before the exclusions against nvdw (ier 51), and the
// table lengths l2cnt.  FILL: each replica counts its entries again, and
// writes them after those of the replicas before it, in candidate order.
template <bool FILL>
__global__ void __launch_bounds__(256) k_cmm_l2(CmmArgs A) {
#pragma clang fp contract(off)
    __shared__ int s_run[CMM_MAXRUN], s_off[CMM_MAXRUN + 1], s_vi[256], s_c[4][256];
    __shared__ double s_x[CMM_L2CH], s_y[CMM_L2CH], s_z[CMM_L2CH];
    __shared__ int s_f[CMM_L2CH], s_a[CMM_L2CH], s_m[CMM_L2CH], s_p[CMM_L2CH];
    const int cell = blockIdx.x, nf = A.nf, nv = A.nv1, t = threadIdx.x;
    if (A.err[0] & (FILL ? 3 : 1)) return;
    int nvi = 0;
    for (int f = 0; f < nf; ++f) nvi += A.kcnt[(cell * nf + f) * 2 + 1];
    if (nvi == 0) return;
    const SysP &P = A.sp[0];
    const int cx = cell % nv, cy = (cell / nv) % nv, cz = cell / (nv * nv);
    const int nslot = cmm_shell_runs(A, cx, cy, cz, A.idelay, s_run, s_off);
    const int J = s_off[nslot];
    for (int i0 = 0; i0 < nvi; i0 += 256) {
        if (t == 0) {
            int w = 0, skip = i0;
            for (int f = 0; f < nf && w < 256; ++f) {
                const int q = (cell * nf + f) * 2 + 1;
                for (int u = 0; u < A.kcnt[q] && w < 256; ++u) {
                    if (skip > 0) { --skip; continue; }
                    s_vi[w++] = A.kstart[q] + u;
                }
            }
        }
        __syncthreads();
        const int nic = min(256, nvi - i0);
        const int R = 256 / nic, r = t / nic, li = t - r * nic;
        const bool act = r < R;
        const int p = s_vi[act ? li : 0];
        const double xi = A.sx0[3 * p], yi = A.sx0[3 * p + 1], zi = A.sx0[3 * p + 2];
        const int fi = (A.sflag[p] >> 13) & 1, ai = A.sorig[p], moli = A.smol[p];
        const unsigned long long mask = A.xmask[ai];
        const bool wide = A.xwide[ai];
        int nvw[3] = {0, 0, 0}, listed = 0, w = 0;
        for (int pass = 0; pass < (FILL ? 2 : 1); ++pass) {
            if (FILL && pass == 1) {   // this replica's entries follow those of replicas 0 .. r-1
                s_c[0][t] = listed;
                __syncthreads();
                w = A.l2ptr[p];
                for (int rr = 0; rr < r; ++rr) w += s_c[0][rr * nic + li];
            }
            for (int q0 = 0; q0 < J; q0 += CMM_L2CH) {
                const int nq = min(CMM_L2CH, J - q0);
                __syncthreads();
                for (int u = t; u < nq; u += 256) {
                    const int g = q0 + u;
                    int lo = 0, hi = nslot - 1;   // the last slot with s_off <= g holds g
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (s_off[mid] <= g) lo = mid;
                        else hi = mid - 1;
                    }
                    const int vj = s_run[lo] + (g - s_off[lo]);
                    const double4 c = A.vx0[vj];
                    s_x[u] = c.x; s_y[u] = c.y; s_z[u] = c.z;
                    s_f[u] = A.vf[vj]; s_a[u] = A.va[vj]; s_m[u] = A.vm[vj]; s_p[u] = A.vp[vj];
                }
                __syncthreads();
                if (act) {
                    for (int u = r; u < nq; u += R) {
                        const int fj = s_f[u];
                        if (!((fj >> 12) & 1)) continue;
                        const double d0 = xi - s_x[u], d1 = yi - s_y[u], d2 = zi - s_z[u];
                        if (!(d0 * d0 + d1 * d1 + d2 * d2 < A.rlim2)) continue;
                        if (!FILL) nvw[cmm_cls(P, fi, (fj >> 13) & 1)]++;
                        if (cmm_excl(A, ai, s_a[u], mask, wide, moli, s_m[u])) continue;
                        if (FILL && pass == 1) A.l2idx[w++] = s_p[u];
                        else ++listed;
                    }
                }
            }
        }
        if (!FILL) {
            __syncthreads();
            s_c[0][t] = act ? listed : 0;
            s_c[1][t] = act ? nvw[0] : 0;
            s_c[2][t] = act ? nvw[1] : 0;
            s_c[3][t] = act ? nvw[2] : 0;
            __syncthreads();
            if (act && r == 0) {
                int c[4] = {0, 0, 0, 0};
                for (int rr = 0; rr < R; ++rr)
                    for (int m = 0; m < 4; ++m) c[m] += s_c[m][rr * nic + li];
                A.l2cnt[p] = c[0];
                if (c[1] > A.capvw || c[2] > A.capvw || c[3] > A.capvw) atomicOr(A.err, 2);
            }
        }
        __syncthreads();
    }
}

// The sorted positions of the atoms with a second-nearest LJ table, in order.
__global__ void k_cmm_vflag(CmmArgs A) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < A.n) A.vflag[p] = A.l2cnt[p] > 0 ? 1 : 0;
}

__global__ void k_cmm_vfill(CmmArgs A) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < A.n && A.l2cnt[p] > 0) A.vlist[A.vpos[p]] = p;
}

// LJ of the second-nearest tables, per step: one wavefront per atom with a
// table, lanes over its entries, a fixed butterfly sum, added to the atom's
// near-field sums (enear's third table).
template <bool MULTI, bool E>
__global__ void __launch_bounds__(256) k_cmm_lj2(CmmArgs A) {
    using Acc = CmmAcc<MULTI>;
    constexpr int NV = Acc::NV, NC = Acc::NC;
    const int w = (int)((blockIdx.x * 256 + threadIdx.x) >> 6), lane = threadIdx.x & 63;
    if (w >= A.vpos[A.n]) return;
    const SysP &P = A.sp[0];
    const int p = A.vlist[w];
    const double4 xi = A.sx[p];
    const int fi = A.sflag[p];
    const int ti = fi & 0xfff, ci = (fi >> 13) & 1, nt = A.ntyp;
    Acc acc;
    acc.zero();
    const int b = A.l2ptr[p], e = b + A.l2cnt[p];
    for (int k = b + lane; k < e; k += 64) {
        const int j = A.l2idx[k];
        const double4 xj = A.sx[j];
        const int fj = A.sflag[j];
        const double dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        const double2 c = A.lj[(fj & 0xfff) * nt + ti];
        const double ri = cmm_rsqrt(dx * dx + dy * dy + dz * dz);
        const double d2i = ri * ri;
        const double d6i = d2i * d2i * d2i;
        const double v1 = c.x * d6i, v2 = c.y * d6i * d6i;
        const double cc = (-6.0 * v1 + 12.0 * v2) * d2i;
        acc.add(MULTI ? cmm_cls(P, ci, (fj >> 13) & 1) : 0, dx, dy, dz, cc, 0.0, v2 - v1);
    }
#pragma unroll
    for (int m = 0; m < NV; ++m)
        for (int o = 32; o > 0; o >>= 1) acc.v[m] = acc.v[m] + __shfl_xor(acc.v[m], o, 64);
    if (lane == 0) {
        double *G = A.gnb + (size_t)p * 9;
        for (int k = 0; k < NC; ++k)
            for (int d = 0; d < 3; ++d) G[3 * k + d] = G[3 * k + d] + acc.v[3 * k + d];
        if (E) {
            double *En = A.enb + (size_t)p * 6;
            for (int k = 0; k < NC; ++k) En[3 + k] = En[3 + k] + acc.v[4 * NC + k];
        }
    }
}

#define CMM_FCH 128      // far-cell records staged per pass of k_cmm_far
#define CMM_MAXFRUN 1024
#define CMM_FP 4         // blocks of k_cmm_far per group, each over a quarter of the records

// Far field (compute_particle_multipole, cmmdo.f90:164-288): for every atom,
// per source fragment and level, the cells of that level in the far list of
// its ancestor (mk_cm: the children of the ancestor's parent's neighbours
// less the ancestor's own neighbours; all cells less the neighbours at the
// top level) that hold atoms of the fragment.
//
// Work is grouped by parent: the 8 level-1 cells of one level-2 cell (one
// level-1 cell when there is a single level) share the far lists of levels
// 2 and up, and their level-1 lists are the children of the parent's
// neighbours, each less its own neighbours (a test per atom).  Block
// (group, part) stages the part-th quarter of the group's records through
// LDS; thread t takes atom t % nic of the group and, as replica t / nic of
// R = 256 / nic, every R-th record; replicas are summed in replica order into
// the part's partial sums, which k_cmm_farsum adds in part order.
template <bool MULTI>
__global__ void __launch_bounds__(CMM_NT) k_cmm_far(CmmArgs A) {
    constexpr int NC = MULTI ? 3 : 1, NV = 4 * NC;
    __shared__ double s_rec[CMM_FCH * 16];
    __shared__ int s_meta[CMM_FCH];
    __shared__ int s_base[CMM_MAXFRUN], s_off[CMM_MAXFRUN + 1], s_inf[CMM_MAXFRUN];
    __shared__ int s_irun[4], s_ioff[5];
    __shared__ double s_red[NV * CMM_NT];
    const int nf = A.nf, nv1 = A.nv1, t = threadIdx.x;
    const int g2 = A.nlev > 1 ? 2 : 1, ng = nv1 / g2;
    const int grp = blockIdx.x / CMM_FP, part = blockIdx.x % CMM_FP;
    const int gx = grp % ng, gy = (grp / ng) % ng, gz = grp / (ng * ng);
    // the group's atoms: g2 x g2 rows of g2 x-adjacent level-1 cells
    if (t < g2 * g2) {
        const int z = g2 * gz + t / g2, y = g2 * gy + t % g2;
        const int b = A.kstart[cmm_lin(nv1, g2 * gx, y, z) * nf * 2];
        s_irun[t] = b;
        s_ioff[t] = A.kstart[(cmm_lin(nv1, g2 * gx + g2 - 1, y, z) + 1) * nf * 2] - b;
    }
    __syncthreads();
    if (t == 0) {
        int off = 0;
        for (int k = 0; k < g2 * g2; ++k) {
            const int len = s_ioff[k];
            s_ioff[k] = off;
            off += len;
        }
        s_ioff[g2 * g2] = off;
    }
    __syncthreads();
    const int NI = s_ioff[g2 * g2];
    if (NI == 0) return;
    const SysP &P = A.sp[0];
    // record slots ((kk, level, row jz, row jy), 2 x-runs each); level 0
    // takes the whole candidate row, the atoms drop their own neighbours
    const int nrow = nf * A.nlev * 36, nslot = 2 * nrow;
    for (int q = t; q < nrow; q += CMM_NT) {
        const int kk = q / (A.nlev * 36), l = (q / 36) % A.nlev, jzi = (q % 36) / 6, jyi = q % 6;
        // level 0: the candidates of the group's cells; level l >= 1: the
        // far list of their common ancestor
        const int ax = (g2 * gx) >> l, ay = (g2 * gy) >> l, az = (g2 * gz) >> l;
        const int mm = A.nv[l];
        int lo[3], hi[3];
        const int ii[3] = {ax, ay, az};
        for (int d = 0; d < 3; ++d) {
            if (l == A.nlev - 1) {
                lo[d] = 0; hi[d] = mm - 1;
            } else {
                const int m1 = A.nv[l + 1], pj = ii[d] >> 1;
                lo[d] = max(pj - 1, 0) * 2;
                hi[d] = min(pj + 1, m1 - 1) * 2 + 1;
            }
        }
        const int jz = lo[2] + jzi, jy = lo[1] + jyi;
        int xa[2] = {0, 0}, xb[2] = {-1, -1};
        if (jz <= hi[2] && jy <= hi[1]) {
            if (l > 0 && abs(jz - az) <= 1 && abs(jy - ay) <= 1) {
                xa[0] = lo[0]; xb[0] = ax - 2;
                xa[1] = ax + 2; xb[1] = hi[0];
            } else {
                xa[0] = lo[0]; xb[0] = hi[0];
            }
        }
        for (int h = 0; h < 2; ++h) {
            const bool on = xa[h] <= xb[h];
            s_base[2 * q + h] = on ? A.pcoff[l] + cmm_lin(mm, xa[h], jy, jz) * nf + kk : 0;
            s_off[2 * q + h] = on ? xb[h] - xa[h] + 1 : 0;
            // kk, level 0 or not, first x, y, z of the run (7 bits each)
            s_inf[2 * q + h] = kk | (l == 0 ? 2 : 0) | (xa[h] << 2) | (jy << 9) | (jz << 16);
        }
    }
    __syncthreads();
    cmm_block_scan(s_off, nslot);
    const int J = s_off[nslot];
    const int pb = (int)((long long)J * part / CMM_FP), pe = (int)((long long)J * (part + 1) / CMM_FP);
    for (int i0 = 0; i0 < NI; i0 += CMM_NT) {
        const int nic = min(CMM_NT, NI - i0);
        const int R = CMM_NT / nic, r = t / nic, li = t - r * nic;
        const bool act = r < R;
        int p = 0;
        if (act) {
            const int g = i0 + li;
            int k = 0;
            while (g >= s_ioff[k + 1]) ++k;
            p = s_irun[k] + (g - s_ioff[k]);
        }
        const double4 xi = act ? A.sx[p] : make_double4(0.0, 0.0, 0.0, 0.0);
        const int k = act ? (A.sflag[p] >> 13) & 1 : 0;
        const int ca = act ? A.key[A.sorig[p]] / (2 * nf) : 0;
        const int cx = ca % nv1, cy = (ca / nv1) % nv1, cz = ca / (nv1 * nv1);
        double G[NV];
#pragma unroll
        for (int m = 0; m < NV; ++m) G[m] = 0.0;
        for (int q0 = pb; q0 < pe; q0 += CMM_FCH) {
            const int nq = min(CMM_FCH, pe - q0);
            __syncthreads();
            for (int u = t; u < nq * 16; u += CMM_NT) {
                const int slot = u >> 4, m = u & 15, g = q0 + slot;
                int lo = 0, hi = nslot - 1;   // the last slot with s_off <= g holds g
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (s_off[mid] <= g) lo = mid;
                    else hi = mid - 1;
                }
                const int dx = g - s_off[lo];
                const int rec = s_base[lo] + dx * nf;
                s_rec[u] = A.pot[(size_t)rec * 16 + m];
                if (m == 0) s_meta[slot] = (A.npc[rec] != 0) ? 1 | (s_inf[lo] + (dx << 2)) << 1 : 0;
            }
            __syncthreads();
            if (act) {
                for (int u = r; u < nq; u += R) {
                    const int meta = s_meta[u];
                    if (!meta) continue;
                    if ((meta >> 2) & 1) {   // level 1: not the atom's own neighbours
                        const int jx = (meta >> 3) & 127, jy = (meta >> 10) & 127, jz = (meta >> 17) & 127;
                        if (abs(jx - cx) <= 1 && abs(jy - cy) <= 1 && abs(jz - cz) <= 1) continue;
                    }
                    const double *c = s_rec + u * 16, *q = c + 3;
                    const double dx = xi.x - c[0], dy = xi.y - c[1], dz = xi.z - c[2];
                    const double dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
                    const double di = cmm_rsqrt(dx2 + dy2 + dz2);
                    const double d2i = di * di, d3i = di * d2i, d5i = d3i * d2i;
                    const double dxy = dx * dy, dyz = dy * dz, dzx = dz * dx;
                    const double aa1 = q[1] * dx + q[2] * dy + q[3] * dz;
                    const double aa2 = q[4] * dx2 + q[5] * dy2 + q[6] * dz2 + q[10] * dxy + q[11] * dyz + q[12] * dzx;
                    const double e = q[0] * di + aa1 * d3i + aa2 * d5i;
                    const double bbx = (q[0] * dx - q[1]) * d3i, bby = (q[0] * dy - q[2]) * d3i,
                                 bbz = (q[0] * dz - q[3]) * d3i;
                    const double cc2x = q[7] * dx + q[10] * dy + q[12] * dz;
                    const double cc2y = q[10] * dx + q[8] * dy + q[11] * dz;
                    const double cc2z = q[12] * dx + q[11] * dy + q[9] * dz;
                    const double cc4 = aa1 * 3.0 + aa2 * d2i * 5.0;
                    const int ncls = MULTI ? cmm_cls(P, (meta >> 1) & 1, k) : 0;
#pragma unroll
                    for (int m = 0; m < NC; ++m) {
                        const bool on = !MULTI || m == ncls;
                        G[4 * m] = G[4 * m] + (on ? -bbx + (cc2x - cc4 * dx) * d5i : 0.0);
                        G[4 * m + 1] = G[4 * m + 1] + (on ? -bby + (cc2y - cc4 * dy) * d5i : 0.0);
                        G[4 * m + 2] = G[4 * m + 2] + (on ? -bbz + (cc2z - cc4 * dz) * d5i : 0.0);
                        G[4 * m + 3] = G[4 * m + 3] + (on ? e : 0.0);
                    }
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < NV; ++m) s_red[m * CMM_NT + t] = G[m];
        __syncthreads();
        if (act && r == 0) {
            double tot[NV];
#pragma unroll
            for (int m = 0; m < NV; ++m) tot[m] = 0.0;
            for (int rr = 0; rr < R; ++rr)
#pragma unroll
                for (int m = 0; m < NV; ++m) tot[m] = tot[m] + s_red[m * CMM_NT + rr * nic + li];
            double *W = A.gwkp + ((size_t)part * A.n + A.sorig[p]) * 12;
            for (int m = 0; m < 12; ++m) W[m] = m < NV ? tot[m] : 0.0;
        }
    }
}

// The parts' sums in part order, scaled by chgmod2 = ccoef q (gradient) and
// ccoef q (energy), into the far field kept between far-field steps.
__global__ void k_cmm_farsum(CmmArgs A) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= A.n) return;
    double t[12];
    for (int m = 0; m < 12; ++m) t[m] = 0.0;
    for (int part = 0; part < CMM_FP; ++part) {
        const double *W = A.gwkp + ((size_t)part * A.n + a) * 12;
        for (int m = 0; m < 12; ++m) t[m] = t[m] + W[m];
    }
    const double qa = A.q[a], q2 = A.ccoef * qa;
    for (int m = 0; m < 3; ++m) {
        for (int d = 0; d < 3; ++d) A.gwk[((size_t)a * 3 + m) * 3 + d] = t[4 * m + d] * q2;
        A.efar[(size_t)a * 3 + m] = qa * A.ccoef * t[4 * m + 3];
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
