// Harmonic bond and angle terms of md_run's dnergv_PB (dnergv.f90:62-193),
// in FP64 with the reference's operation order (no FMA contraction).
//
// Terms are stored in original atom numbering:
//   bond  t: bnd[t]  = {i, j, class, 0},  bndp[t] = {fc, r0}
//   angle t: ang[t]  = {i, j, k, class},  angp[t] = {fc, theta0}
// The gradient is gathered per atom: bat[] lists the atoms that carry terms,
// bref[bptr[u] .. bptr[u+1]) the (term, role) codes of atom bat[u], so every
// atom sums its own contributions in a fixed order and no atomics are needed.
#pragma once
#include "jp_kernels.hpp"

#define JP_BEPS 1.0e-10   // PHYCNS.f90 eps: floor of sin(theta) in dneang_PB

// term code: bits 0-1 role (bond 0/1, angle 0/1/2), bit 2 angle flag, bits 3+ index
__host__ __device__ inline int bref_code(int idx, bool angle, int role) { return (idx << 3) | (angle ? 4 : 0) | role; }

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

struct BondedArgs {
    int nbat;
    const int *bat, *bptr, *bref;
    const int4 *bnd, *ang;
    const double2 *bndp, *angp;
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
        const int t = code >> 3, role = code & 3;
        double c[3];
        int cls;
        if (code & 4) {
            const int4 an = B.ang[t];
            double w1[3], w3[3];
            angle_term(P, x, an, B.angp[t], w1, w3);
            cls = an.w;
            for (int k = 0; k < 3; ++k) c[k] = role == 0 ? w1[k] : (role == 2 ? w3[k] : -(w1[k] + w3[k]));
        } else {
            const int4 bd = B.bnd[t];
            double w[3];
            bond_term(P, x, bd, B.bndp[t], w);
            cls = bd.z;
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

// Energy of every term (bonds first, then angles), summed per class on the
// host in term order, as dnebnd_PB / dneang_PB accumulate them.
__global__ void k_bonded_energy(int nbond, int nangle, BondedArgs B, const int *__restrict__ atom_sys,
                                const SysP *__restrict__ sp, const double *__restrict__ x, double *__restrict__ e) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nbond) {
        const int4 bd = B.bnd[t];
        double w[3];
        e[t] = bond_term(sp[atom_sys[bd.x]], x, bd, B.bndp[t], w);
    } else if (t < nbond + nangle) {
        const int4 an = B.ang[t - nbond];
        double w1[3], w3[3];
        e[t] = angle_term(sp[atom_sys[an.x]], x, an, B.angp[t - nbond], w1, w3);
    }
}
