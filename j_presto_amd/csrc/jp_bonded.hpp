// Bonded terms of md_run's dnergv_PB (dnergv.f90:29-516): harmonic bonds and
// angles, torsions and impropers (dnetor_PB), 1-4 LJ and electrostatics
// (dne14_PB), in FP64 with the reference's operation order (no FMA
// contraction).
//
// Terms are stored in original atom numbering:
//   bond     t: bnd[t] = {i, j, class, 0},     bndp[t] = {fc, r0}
//   angle    t: ang[t] = {i, j, k, class},     angp[t] = {fc, theta0}
//   torsion  t: tor[t] = {i, j, k, l},         torp[t] = {fc, periodicity, phase, division},
//               torc[t] = class | improper << 8 (impropers follow the torsions)
//   1-4 pair t: p14[t] = {i, j, class, 0},     p14p[t] = {LJ scale, scale * ccoef * q_i q_j}
// The gradient is gathered per atom: bat[] lists the atoms that carry terms,
// bref[bptr[u] .. bptr[u+1]) the (term, role) codes of atom bat[u], so every
// atom sums its own contributions in a fixed order and no atomics are needed.
#pragma once
#include "jp_kernels.hpp"

#define JP_BEPS 1.0e-10   // PHYCNS.f90 eps: floor of sin(theta) in dneang_PB

// term code: bits 0-1 role (atom of the term), bits 2-3 kind, bits 4+ index
#define JP_BK_BOND 0
#define JP_BK_ANGLE 1
#define JP_BK_TORS 2
#define JP_BK_P14 3
__host__ __device__ inline int bref_code(int idx, int kind, int role) { return (idx << 4) | (kind << 2) | role; }

__device__ __forceinline__ void minimg_d(const SysP &P, double *d) {
#pragma clang fp contract(off)
    for (int k = 0; k < 3; ++k) d[k] = d[k] - P.L[k] * anint_d(d[k] * P.invL[k]);
}

// dnebnd_PB (dnergv.f90:93-111): energy, and work = 2 fc (r - r0) d / r with
// d = x_i - x_j (grad_i += work, grad_j -= work).
__device__ inline double bond_term(const SysP &P, const double *__restrict__ x, int4 b, double2 p, double *work) {
#pragma clang fp contract(off)
    double d[3];
    for (int k = 0; k < 3; ++k) d[k] = x[3 * b.x + k] - x[3 * b.y + k];
    minimg_d(P, d);
    const double r12 = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double dif = r12 - p.y;
    double coef = p.x * dif;
    const double e = coef * dif;
    coef = 2.0 * coef / r12;
    for (int k = 0; k < 3; ++k) work[k] = coef * d[k];
    return e;
}

// dneang_PB (dnergv.f90:142-182): energy, w1 (atom 1) and w3 (atom 3);
// atom 2 receives -(w1 + w3).
__device__ inline double angle_term(const SysP &P, const double *__restrict__ x, int4 a, double2 p, double *w1,
                                    double *w3) {
#pragma clang fp contract(off)
    double d12[3], d32[3];
    for (int k = 0; k < 3; ++k) {
        d12[k] = x[3 * a.x + k] - x[3 * a.y + k];
        d32[k] = x[3 * a.z + k] - x[3 * a.y + k];
    }
    minimg_d(P, d12);
    minimg_d(P, d32);
    double r12 = d12[0] * d12[0] + d12[1] * d12[1] + d12[2] * d12[2];
    double r32 = d32[0] * d32[0] + d32[1] * d32[1] + d32[2] * d32[2];
    const double r12r32 = sqrt(1.0 / (r12 * r32));
    double cost = d12[0] * d32[0] + d12[1] * d32[1] + d12[2] * d32[2];
    r12 = 1.0 / r12;
    r32 = 1.0 / r32;
    cost = fmin(cost * r12r32, 1.0);
    cost = fmax(cost, -1.0);
    const double theta = acos(cost);
    const double dif = theta - p.y;
    double coef = p.x * dif;
    const double e = coef * dif;
    const double sint = fmax(sin(theta), JP_BEPS);
    coef = -2.0 * coef / sint;
    double rtmp = cost * r12;
    for (int k = 0; k < 3; ++k) w1[k] = coef * (d32[k] * r12r32 - d12[k] * rtmp);
    rtmp = cost * r32;
    for (int k = 0; k < 3; ++k) w3[k] = coef * (d12[k] * r12r32 - d32[k] * rtmp);
    return e;
}

// dnetor_PB (dnergv.f90:314-426): energy and work[9] (atoms 1, 2 and 4; atom 3
// receives -(w1 + w2 + w4)).  Returns false (no term) when a bond angle is 0
// or pi, as the reference skips it.
__device__ inline bool torsion_term(const SysP &P, const double *__restrict__ x, int4 a, double4 p, double &e,
                                    double *work) {
#pragma clang fp contract(off)
    double d21[3], d32[3], d43[3], p12[3], p23[3], p123[3];
    for (int k = 0; k < 3; ++k) {
        d21[k] = x[3 * a.y + k] - x[3 * a.x + k];
        d32[k] = x[3 * a.z + k] - x[3 * a.y + k];
        d43[k] = x[3 * a.w + k] - x[3 * a.z + k];
    }
    minimg_d(P, d21);
    minimg_d(P, d32);
    minimg_d(P, d43);
    p12[0] = d21[1] * d32[2] - d32[1] * d21[2];
    p12[1] = d21[2] * d32[0] - d32[2] * d21[0];
    p12[2] = d21[0] * d32[1] - d32[0] * d21[1];
    p23[0] = d32[1] * d43[2] - d43[1] * d32[2];
    p23[1] = d32[2] * d43[0] - d43[2] * d32[0];
    p23[2] = d32[0] * d43[1] - d43[0] * d32[1];
    double r12 = p12[0] * p12[0] + p12[1] * p12[1] + p12[2] * p12[2];
    double r23 = p23[0] * p23[0] + p23[1] * p23[1] + p23[2] * p23[2];
    double r12r23 = sqrt(r12 * r23);
    if (!(r12r23 > 0.0)) return false;
    r12r23 = 1.0 / r12r23;
    const double s1223 = p12[0] * p23[0] + p12[1] * p23[1] + p12[2] * p23[2];
    double cosp = s1223 * r12r23;
    cosp = cosp > 1.0 ? 1.0 : (cosp < -1.0 ? -1.0 : cosp);
    p123[0] = p12[1] * p23[2] - p23[1] * p12[2];
    p123[1] = p12[2] * p23[0] - p23[2] * p12[0];
    p123[2] = p12[0] * p23[1] - p23[0] * p12[1];
    const double s32123 = d32[0] * p123[0] + d32[1] * p123[1] + d32[2] * p123[2];
    double phi = acos(cosp);
    if (s32123 < 0.0) phi = 2.0 * 3.141592653589793 - phi;   // PHYCNS.f90:11 pi
    const double phimod = p.y * phi + p.z;
    const double div = 1.0 / p.w;
    e = p.x * (1.0 + cos(phimod)) * div;
    const double sinp = sin(phi);
    double wrk0;
    switch ((int)anint_d(p.y)) {   // periodicity 1-6 (checked on the host)
        case 1: wrk0 = 1.0; break;
        case 2: wrk0 = 2.0 * cosp; break;
        case 3: wrk0 = -4.0 * sinp * sinp + 3.0; break;
        case 4: wrk0 = (-8.0 * sinp * sinp + 4.0) * cosp; break;
        case 5: wrk0 = (16.0 * sinp * sinp - 20.0) * sinp * sinp + 5.0; break;
        default: wrk0 = cosp * (6.0 + 32.0 * cosp * cosp * (cosp * cosp - 1.0)); break;
    }
    wrk0 = cos(p.z) * p.y * p.x * div * wrk0 * r12r23;
    const double wrk1 = d21[0] + d32[0], wrk2 = d21[1] + d32[1], wrk3 = d21[2] + d32[2];
    r12 = s1223 / r12;
    r23 = s1223 / r23;
    work[0] = d32[2] * p23[1] - d32[1] * p23[2] - r12 * (d32[2] * p12[1] - d32[1] * p12[2]);
    work[1] = d32[0] * p23[2] - d32[2] * p23[0] - r12 * (d32[0] * p12[2] - d32[2] * p12[0]);
    work[2] = d32[1] * p23[0] - d32[0] * p23[1] - r12 * (d32[1] * p12[0] - d32[0] * p12[1]);
    work[3] = -wrk3 * p23[1] + d43[2] * p12[1] + wrk2 * p23[2] - d43[1] * p12[2] -
              r12 * (-wrk3 * p12[1] + wrk2 * p12[2]) - r23 * (d43[2] * p23[1] - d43[1] * p23[2]);
    work[4] = wrk3 * p23[0] - d43[2] * p12[0] - wrk1 * p23[2] + d43[0] * p12[2] -
              r12 * (wrk3 * p12[0] - wrk1 * p12[2]) - r23 * (d43[0] * p23[2] - d43[2] * p23[0]);
    work[5] = -wrk2 * p23[0] + d43[1] * p12[0] + wrk1 * p23[1] - d43[0] * p12[1] -
              r12 * (-wrk2 * p12[0] + wrk1 * p12[1]) - r23 * (d43[1] * p23[0] - d43[0] * p23[1]);
    work[6] = d32[2] * p12[1] - d32[1] * p12[2] - r23 * (d32[2] * p23[1] - d32[1] * p23[2]);
    work[7] = d32[0] * p12[2] - d32[2] * p12[0] - r23 * (d32[0] * p23[2] - d32[2] * p23[0]);
    work[8] = d32[1] * p12[0] - d32[0] * p12[1] - r23 * (d32[1] * p23[0] - d32[0] * p23[1]);
    for (int k = 0; k < 9; ++k) work[k] = work[k] * wrk0;
    return true;
}

// dne14_PB (dnergv.f90:473-511): LJ and electrostatic energies and work (atom
// 1 += work, atom 2 -= work).  c612 = {C6, C12} of the pair's types.
__device__ inline void pair14_term(const SysP &P, const double *__restrict__ x, int4 a, double2 p, double2 c612,
                                   double &ev, double &ee, double *work) {
#pragma clang fp contract(off)
    double d14[3];
    for (int k = 0; k < 3; ++k) d14[k] = x[3 * a.x + k] - x[3 * a.y + k];
    minimg_d(P, d14);
    double r14 = d14[0] * d14[0] + d14[1] * d14[1] + d14[2] * d14[2];
    r14 = 1.0 / r14;
    const double r141 = sqrt(r14);
    const double r146 = r14 * r14 * r14;
    const double r1412 = r146 * r146;
    const double co6 = c612.x * r146, co12 = c612.y * r1412;
    ev = p.x * (-co6 + co12);
    double coef = r14 * p.x * (-2.0 * co12 + co6) * 6.0;
    for (int k = 0; k < 3; ++k) work[k] = coef * d14[k];
    coef = p.y * r141;
    ee = coef;
    coef = coef * (-r14);
    for (int k = 0; k < 3; ++k) work[k] = work[k] + coef * d14[k];
}

struct BondedArgs {
    int nbat;
    const int *bat, *bptr, *bref;
    const int4 *bnd, *ang, *tor, *p14;
    const double2 *bndp, *angp, *p14p;
    const double4 *torp;
    const int *torc;
    const int *tflag;       // per atom: LJ type in bits 0-11
    const double2 *lj64;    // C6, C12 per type pair (system table at SysP::lj0)
};

// Bonded gradient of every atom in bat[], scaled by its term's class factor
// (efac[3 * system + class]).  gout == nullptr: added to the atom's slot of
// gslot (after the pair kernel wrote it); else stored to gout in original
// order (gout pre-zeroed by the caller).
__global__ void __launch_bounds__(128)
    k_bonded(BondedArgs B, const int *__restrict__ atom_sys, const SysP *__restrict__ sp, const double *__restrict__ x,
             const double *__restrict__ efac, const int *__restrict__ inum, double *__restrict__ gslot,
             double *__restrict__ gout) {
#pragma clang fp contract(off)
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= B.nbat) return;
    const int a = B.bat[u];
    const int s = atom_sys[a];
    const SysP &P = sp[s];
    double g[3] = {0.0, 0.0, 0.0};
    for (int q = B.bptr[u]; q < B.bptr[u + 1]; ++q) {
        const int code = B.bref[q];
        const int t = code >> 4, kind = (code >> 2) & 3, role = code & 3;
        double c[3];
        int cls;
        if (kind == JP_BK_ANGLE) {
            const int4 an = B.ang[t];
            double w1[3], w3[3];
            angle_term(P, x, an, B.angp[t], w1, w3);
            cls = an.w;
            for (int k = 0; k < 3; ++k) c[k] = role == 0 ? w1[k] : (role == 2 ? w3[k] : -(w1[k] + w3[k]));
        } else if (kind == JP_BK_BOND) {
            const int4 bd = B.bnd[t];
            double w[3];
            bond_term(P, x, bd, B.bndp[t], w);
            cls = bd.z;
            for (int k = 0; k < 3; ++k) c[k] = role == 0 ? w[k] : -w[k];
        } else if (kind == JP_BK_TORS) {
            double w[9], e;
            cls = B.torc[t] & 0xff;
            if (torsion_term(P, x, B.tor[t], B.torp[t], e, w)) {
                for (int k = 0; k < 3; ++k)
                    c[k] = role == 0 ? w[k] : (role == 1 ? w[3 + k] : (role == 3 ? w[6 + k] : -(w[k] + w[3 + k] + w[6 + k])));
            } else {
                c[0] = c[1] = c[2] = 0.0;
            }
        } else {
            const int4 pr = B.p14[t];
            const int ti = B.tflag[pr.x] & 0xfff, tj = B.tflag[pr.y] & 0xfff;
            double w[3], ev, ee;
            pair14_term(P, x, pr, B.p14p[t], B.lj64[P.lj0 + ti * P.ntyp + tj], ev, ee, w);
            cls = pr.z;
            for (int k = 0; k < 3; ++k) c[k] = role == 0 ? w[k] : -w[k];
        }
        const double f = efac[3 * s + cls];
        for (int k = 0; k < 3; ++k) g[k] = g[k] + c[k] * f;
    }
    if (gout) {
        for (int k = 0; k < 3; ++k) gout[3 * a + k] = g[k];
    } else {
        const size_t sl = inum[a];
        for (int k = 0; k < 3; ++k) gslot[3 * sl + k] = gslot[3 * sl + k] + g[k];
    }
}

// Energy of every term: bonds, angles, torsions (impropers after them), then
// the LJ and the electrostatic energy of each 1-4 pair (two entries); summed
// per class and energy slot on the host in term order, as the reference
// routines accumulate them.
__global__ void k_bonded_energy(int nbond, int nangle, int ntors, int n14, BondedArgs B,
                                const int *__restrict__ atom_sys, const SysP *__restrict__ sp,
                                const double *__restrict__ x, double *__restrict__ e) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nbond) {
        const int4 bd = B.bnd[t];
        double w[3];
        e[t] = bond_term(sp[atom_sys[bd.x]], x, bd, B.bndp[t], w);
        return;
    }
    t -= nbond;
    if (t < nangle) {
        const int4 an = B.ang[t];
        double w1[3], w3[3];
        e[nbond + t] = angle_term(sp[atom_sys[an.x]], x, an, B.angp[t], w1, w3);
        return;
    }
    t -= nangle;
    if (t < ntors) {
        const int4 a = B.tor[t];
        double w[9], en = 0.0;
        if (!torsion_term(sp[atom_sys[a.x]], x, a, B.torp[t], en, w)) en = 0.0;
        e[nbond + nangle + t] = en;
        return;
    }
    t -= ntors;
    if (t < n14) {
        const int4 pr = B.p14[t];
        const SysP &P = sp[atom_sys[pr.x]];
        const int ti = B.tflag[pr.x] & 0xfff, tj = B.tflag[pr.y] & 0xfff;
        double w[3], ev, ee;
        pair14_term(P, x, pr, B.p14p[t], B.lj64[P.lj0 + ti * P.ntyp + tj], ev, ee, w);
        e[nbond + nangle + ntors + 2 * t] = ev;
        e[nbond + nangle + ntors + 2 * t + 1] = ee;
    }
}
