// Canonical MD (md_run METHOD= CANO, verlethe.f90:79-757 with idfvsl = 3).
//
// One step after the forces:
//   begin   v_b = v, x_o = x, f_c = 0, w = v - (dt/2) c g / m  -> m|w|^2
//   round 1 drift: chi = sqrt(T_set / T_g(w)) (1 without SHAKE),
//           v = (2 chi - 1) v - chi dt c g / m, x += dt v; then as below
//   round r > 1, skipped once the system is done: the end of round r - 1
//           (every group within fcstol T_set: done), else
//           f_c += (x - x_b)/dt^2, chi from w' below, v = (-c g / m + f_c)
//           chi dt + v_b (2 chi - 1), x = x_o + dt v; then as below
//   every round: x_b = x, SHAKE(x; x_o), v = (x - x_o)/dt, w = (v_b + v)/2
//           -> m|w|^2 for the end-of-round test, and the trial velocity of
//           the next round w' = (dt/2)(-c g / m + f_c + (x - x_b)/dt^2) + v_b
//           -> m|w'|^2, so that the next round needs no kernel of its own
//           to reduce them
// Each round is one launch (k_cano_shake); a system without SHAKE groups
// takes only round 1's drift.  The sorted coordinates follow x.  Under
// STOPCE each round is two launches with the centre-of-mass motion taken
// out of the updated velocities in between (k_remotv, verlethe.f90:489-502,
// 705-718).
// T_g = (sum over group g of m|w|^2) 1e7 / (dof_g gascst) (caltmp.f90), with
// one group per system (TEMPCO= NO) or per molecule kind and cluster
// (TEMPCO= YES, caltmp2).  c = 4.184e-4 (md.f90:56), 1/m is md_run's ifxmass.
// Group sums: every block of a kernel that produces m|w|^2 leaves one
// fixed-order partial per group of its system (blocks hold up to 256 atoms,
// or 128 SHAKE units, of one system, counted from the system's first); every
// block of the next kernel sums its system's partials in one fixed order, so
// all blocks see the same temperatures, runs are deterministic and a system
// gives the same result alone or in a batch.  The order
// differs from md_run's serial sum, which moves results in the last bits.
// Expressions keep md_run's operand order.
#pragma once
#include "jp_kernels.hpp"

#define JP_GASCST 8.31451   // PHYCNS.f90:21
#define JP_FDCVCV 4.184e-4  // md.f90:56

#define JP_CANO_MAXG 32   // temperature groups of one system
#define JP_LS 10          // doubles of ALSD state per system (jp_alsd.hpp)

struct CanoArgs {
    int natom, nunit, icslop, nab, nub;   // nab, nub: atom blocks, unit blocks
    int ngrp;                  // temperature groups of all systems
    const int2 *ablk, *ublk;   // per block: [first, end) atoms (<= 256) or units (<= 128) of one system
    double fcstol;
    const int *atom_sys, *inum, *img;
    const int *gid;            // temperature group of each atom
    const int *gsys;           // system of each group
    const int *sg0, *sg1;      // groups of system s: [sg0[s], sg1[s])
    const int *ab0, *ab1;      // atom blocks of system s: [ab0[s], ab1[s])
    const int *ub0, *ub1;      // unit blocks of system s
    const int *sfirst;         // first atom of system s (writes the system's flag)
    const int *shk_on;         // system s has SHAKE groups
    const double *gdof, *temps;
    const double *ttgt;        // per system the step's target temperature: temps, or the heating
                               // loop's ramp (verlethe.f90:419-423, 600-604)
    const double *mass, *imass, *gslot;
    const double *gew;         // METHOD= GE: per system, the weight every gradient read carries (null: 1)
    // ALSD (jp_alsd.hpp; lam null: off): per system the lambda state
    // [JP_LS], its mass, the group whose temperature takes lambda's kinetic
    // term (-1: none), and l_temp_control
    double *lam;
    const double *lmass;
    const int *lgrp;
    int ltc;
    double *x, *v, *xold, *vback, *fcc, *xcb;
    double *pa;                // partials of m|w|^2 of begin: [group][atom block]
    double *pu, *pn;           // of w and w' per round, two buffers by round parity:
                               // [round & 1][group][unit block]
    int *done;                 // per system: the SHAKE/temperature rounds have ended
};

// Groups of the system of atom a0.
__device__ __forceinline__ void block_groups(const CanoArgs &A, int a0, int &g0, int &g1) {
    const int s = A.atom_sys[a0];
    g0 = A.sg0[s];
    g1 = A.sg1[s];
}

// Per group g in [g0, g1): the sum of the values val[q] the threads hold for
// group gq[q] (butterfly per wave, waves in order) into P[g * nblk + block].
// All threads of the block take part.
template <int NQ>
__device__ __forceinline__ void block_partials(const double (&val)[NQ], const int (&gq)[NQ], int g0, int g1,
                                               double *__restrict__ P, int nblk) {
#pragma clang fp contract(off)
    __shared__ double sw[4];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63, nw = blockDim.x >> 6;
    for (int g = g0; g < g1; ++g) {
        double p = 0.0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) p = p + (gq[q] == g ? val[q] : 0.0);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) p = p + __shfl_xor(p, o, 64);
        if (l == 0) sw[w] = p;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int q = 0; q < nw; ++q) t = t + sw[q];
            P[(size_t)g * nblk + blockIdx.x] = t;
        }
        __syncthreads();
    }
}

// ALSD: lambda's kinetic term m_l lv^2 (lv from state slot ls) where group
// g takes it (caltmp.f90:24, 98-99), else 0.
__device__ __forceinline__ double lambda_ke(const CanoArgs &A, int s, int g, int ls) {
#pragma clang fp contract(off)
    if (!A.lam || A.lgrp[s] != g) return 0.0;
    const double lv = A.lam[JP_LS * s + ls];
    return A.lmass[s] * lv * lv;
}

// Temperatures of groups [g0, g1) into sT (caltmp.f90:39-43) from the
// partials of the blocks [bl0[s], bl1[s]) of each group's system: lane l adds
// blocks l, l+64, ... in order, then a butterfly (every block gets the same
// value), plus lambda's kinetic term (ALSD, velocity in state slot ls).  All
// threads of the block take part.
__device__ __forceinline__ void block_temps(const CanoArgs &A, const double *__restrict__ P, int nblk,
                                            const int *__restrict__ bl0, const int *__restrict__ bl1, int g0, int g1,
                                            double *sT, int ls) {
#pragma clang fp contract(off)
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63, nw = blockDim.x >> 6;
    for (int g = g0 + w; g < g1; g += nw) {
        const int s = A.gsys[g];
        double p = 0.0;
        for (int b = bl0[s] + l; b < bl1[s]; b += 64) p = p + P[(size_t)g * nblk + b];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) p = p + __shfl_xor(p, o, 64);
        p = lambda_ke(A, s, g, ls) + p;
        if (l == 0) sT[g - g0] = p * 1.0e7 / (A.gdof[g] * JP_GASCST);
    }
    __syncthreads();
}

// ALSD, round 1 of a step (verlethe.f90:438-445, 477-482, 508): lambda's
// velocity scaled with the temperature T of its group when l_temp_control
// (else a plain kick), then lambda's drift, and the mean velocity the later
// rounds' temperatures use.  One thread per system.
// T: the temperature of lambda's group; tt: the step's target temperature.
__device__ __forceinline__ void alsd_move(const CanoArgs &A, int s, double T, double tt, double dt) {
#pragma clang fp contract(off)
    double *L = A.lam + JP_LS * s;
    const double lf = L[2], ilm = 1.0 / A.lmass[s];
    double lv = L[1];
    if (A.ltc && A.lgrp[s] >= 0) {
        const double itc = 1.0 / T;
        const double chi = sqrt(tt * itc);
        const double rt = chi * dt * JP_FDCVCV, rt2 = 2.0 * chi - 1.0;
        lv = rt2 * lv + rt * lf * ilm;
    } else {
        lv = lv + dt * JP_FDCVCV * lf * ilm;
    }
    L[1] = lv;
    L[0] = L[0] + dt * lv;
    L[6] = (L[5] + lv) * 0.5;
}

// block_temps of two partial arrays in one pass: P into sT, P2 into sT2.
__device__ __forceinline__ void block_temps2(const CanoArgs &A, const double *__restrict__ P,
                                             const double *__restrict__ P2, int nblk, const int *__restrict__ bl0,
                                             const int *__restrict__ bl1, int g0, int g1, double *sT, double *sT2) {
#pragma clang fp contract(off)
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63, nw = blockDim.x >> 6;
    for (int g = g0 + w; g < g1; g += nw) {
        const int s = A.gsys[g];
        double p = 0.0, p2 = 0.0;
        for (int b = bl0[s] + l; b < bl1[s]; b += 64) {
            p = p + P[(size_t)g * nblk + b];
            p2 = p2 + P2[(size_t)g * nblk + b];
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            p = p + __shfl_xor(p, o, 64);
            p2 = p2 + __shfl_xor(p2, o, 64);
        }
        const double ke = lambda_ke(A, s, g, 6);
        p = ke + p;
        p2 = ke + p2;
        if (l == 0) {
            sT[g - g0] = p * 1.0e7 / (A.gdof[g] * JP_GASCST);
            sT2[g - g0] = p2 * 1.0e7 / (A.gdof[g] * JP_GASCST);
        }
    }
    __syncthreads();
}

// Gradient component k of slot sl as md_run uses it: under METHOD= GE
// scaled in place by the multicanonical weight (verlethe.f90:223-233).
__device__ __forceinline__ double cano_grad(const CanoArgs &A, int s, size_t sl, int k) {
#pragma clang fp contract(off)
    const double g = A.gslot[3 * sl + k];
    return A.gew ? g * A.gew[s] : g;
}

__device__ __forceinline__ double cano_mw2(double m, double w0, double w1, double w2) {
#pragma clang fp contract(off)
    return m * (w0 * w0 + w1 * w1 + w2 * w2);
}

__device__ __forceinline__ void cano_sorted(const SysP &P, const CanoArgs &A, int a, const double *xa,
                                            JAtom *__restrict__ ys) {
#pragma clang fp contract(off)
    JAtom *j = ys + A.inum[a];
    j->x = xa[0] - P.L[0] * (double)A.img[3 * a];
    j->y = xa[1] - P.L[1] * (double)A.img[3 * a + 1];
    j->z = xa[2] - P.L[2] * (double)A.img[3 * a + 2];
}

__global__ void __launch_bounds__(256) k_cano_begin(CanoArgs A, const SysP *__restrict__ sp) {
#pragma clang fp contract(off)
    const int2 B = A.ablk[blockIdx.x];
    const int a = B.x + threadIdx.x;
    double val[1] = {0.0};
    int gq[1] = {-1};
    if (a < B.y) {
        const int s = A.atom_sys[a];
        const double dt = sp[s].dt, im = A.imass[a], r = 0.5 * dt * JP_FDCVCV;
        const size_t sl = (size_t)A.inum[a];
        double w[3];
        for (int k = 0; k < 3; ++k) {
            const double vk = A.v[3 * a + k];
            A.vback[3 * a + k] = vk;
            A.xold[3 * a + k] = A.x[3 * a + k];
            A.fcc[3 * a + k] = 0.0;
            w[k] = vk - r * cano_grad(A, s, sl, k) * im;
        }
        val[0] = cano_mw2(A.mass[a], w[0], w[1], w[2]);
        gq[0] = A.gid[a];
        if (a == A.sfirst[s]) A.done[s] = A.shk_on[s] ? 0 : 1;
    }
    int g0, g1;
    block_groups(A, B.x, g0, g1);
    block_partials(val, gq, g0, g1, A.pa, A.nab);
}

// phase 0: the velocity update and the drift; 1: the update alone (v
// written); 2: the drift alone from v (the centre-of-mass motion taken out
// in between, k_remotv).
__global__ void __launch_bounds__(256) k_cano_drift(CanoArgs A, const SysP *__restrict__ sp, JAtom *__restrict__ ys,
                                                    int phase) {
#pragma clang fp contract(off)
    __shared__ double sT[JP_CANO_MAXG];
    const int2 B = A.ablk[blockIdx.x];
    const int a = B.x + threadIdx.x;
    if (phase == 2) {
        if (a >= B.y) return;
        const int s = A.atom_sys[a];
        const SysP &P = sp[s];
        double xa[3];
        for (int k = 0; k < 3; ++k) {
            xa[k] = A.x[3 * a + k] + P.dt * A.v[3 * a + k];
            A.x[3 * a + k] = xa[k];
        }
        if (!A.shk_on[s]) cano_sorted(P, A, a, xa, ys);
        return;
    }
    int g0, g1;
    block_groups(A, B.x, g0, g1);
    block_temps(A, A.pa, A.nab, A.ab0, A.ab1, g0, g1, sT, 3);
    if (a >= B.y) return;
    const int s = A.atom_sys[a];
    const SysP &P = sp[s];
    const double dt = P.dt, im = A.imass[a], ts = A.temps[s], tt = A.ttgt[s];
    const double T = A.shk_on[s] ? sT[A.gid[a] - g0] : ts;   // verlethe.f90:396-415
    if (A.lam && a == A.sfirst[s]) {
        const int lg = A.lgrp[s];
        alsd_move(A, s, (A.shk_on[s] && lg >= 0) ? sT[lg - g0] : ts, tt, dt);
    }
    const double itc = 1.0 / T;
    const double chi = sqrt(tt * itc);
    const double rt = chi * dt * JP_FDCVCV, rt2 = 2.0 * chi - 1.0;
    const size_t sl = (size_t)A.inum[a];
    double xa[3];
    for (int k = 0; k < 3; ++k) {
        const double vk = rt2 * A.v[3 * a + k] - rt * cano_grad(A, s, sl, k) * im;
        A.v[3 * a + k] = vk;
        if (phase == 1) continue;
        xa[k] = A.x[3 * a + k] + dt * vk;
        A.x[3 * a + k] = xa[k];
    }
    if (phase == 0 && !A.shk_on[s]) cano_sorted(P, A, a, xa, ys);
}

// One round per unit (SHAKE group or free atom): round 1 first applies the
// drift, later rounds the end-of-round test of the previous one and its
// velocity update (verlethe.f90:396-404, 601-611, 646-657); then SHAKE as in
// k_step.  A block holds units of one system, so the decisions below are
// uniform over the block, and every block of a system reaches the same one
// from the same fixed-order sums.  PH 0: the whole round; PH 1: up to the
// velocity update (v and the constraint force written); PH 2: the rest from
// v, the centre-of-mass motion taken out in between (k_remotv).
template <int PH, int NQ>
__global__ void __launch_bounds__(128) k_cano_shake(CanoArgs A, const SysP *__restrict__ sp,
                                                    const int4 *__restrict__ unit, const double *__restrict__ s2d,
                                                    const double *__restrict__ s3d, const int4 *__restrict__ s4a,
                                                    const double *__restrict__ s4d, double tol, int maxit,
                                                    int round, JAtom *__restrict__ ys, int *__restrict__ err) {
#pragma clang fp contract(off)
    __shared__ double sT[JP_CANO_MAXG], sU[JP_CANO_MAXG];
    __shared__ int s_run;
    const int2 B = A.ublk[blockIdx.x];
    const int u = B.x + threadIdx.x;
    int g0, g1;
    block_groups(A, unit[B.x].x, g0, g1);
    const int s = A.atom_sys[unit[B.x].x];
    const size_t half = (size_t)A.ngrp * A.nub;
    double *pu_out = A.pu + (size_t)(round & 1) * half, *pn_out = A.pn + (size_t)(round & 1) * half;
    const double *pu_in = A.pu + (size_t)((round - 1) & 1) * half;
    const double *pn_in = A.pn + (size_t)((round - 1) & 1) * half;
    const double ts = A.temps[s], tt = A.ttgt[s];
    // the unit's atom data, loaded before the group sums so that both
    // memory round trips overlap: x, x_o, g, f_c, v_b, and v (round 1) or
    // x_b (later rounds)
    const int4 U = unit[min(u, B.y - 1)];
    const int kind = U.w & 7, prm = U.w >> 3;
    int a[NQ];
    a[0] = U.x; a[1] = U.y; a[2] = U.z;
    if constexpr (NQ == 4) a[3] = kind == 4 ? s4a[prm].w : U.x;
    const bool live = u < B.y;
    double ax[NQ][3], axo[NQ][3], ag[NQ][3], afc[NQ][3], avb[NQ][3], avx[NQ][3], aim[NQ], am[NQ];
    int agid[NQ], asl[NQ], aimg[NQ][3];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int b = (live && q < kind) ? a[q] : a[0];
        const size_t sl = (size_t)A.inum[b];
        asl[q] = (int)sl;
        aim[q] = A.imass[b];
        am[q] = A.mass[b];
        agid[q] = A.gid[b];
        for (int k = 0; k < 3; ++k) aimg[q][k] = A.img[3 * b + k];
        for (int k = 0; k < 3; ++k) {
            const int qq = 3 * b + k;
            ax[q][k] = A.x[qq];
            axo[q][k] = A.xold[qq];
            ag[q][k] = cano_grad(A, s, sl, k);
            afc[q][k] = A.fcc[qq];
            avb[q][k] = A.vback[qq];
            avx[q][k] = (round == 1 || PH == 2) ? A.v[qq] : A.xcb[qq];
        }
    }
    if (PH == 2) {
        if (round > 1) {
            if (threadIdx.x == 0) s_run = !A.done[s];
            __syncthreads();
            if (!s_run) return;   // uniform over the block
        }
    } else if (round == 1) {
        block_temps(A, A.pa, A.nab, A.ab0, A.ab1, g0, g1, sT, 3);
        if (A.lam && threadIdx.x == 0 && blockIdx.x == A.ub0[s]) {
            const int lg = A.lgrp[s];
            alsd_move(A, s, (A.shk_on[s] && lg >= 0) ? sT[lg - g0] : ts, tt, sp[s].dt);
        }
    } else {
        if (threadIdx.x == 0) s_run = !A.done[s];
        __syncthreads();
        if (!s_run) return;   // uniform over the block
        // sU: temperatures of w (end-of-round test), sT: of w' (the update)
        block_temps2(A, pu_in, pn_in, A.nub, A.ub0, A.ub1, g0, g1, sU, sT);
        double tmin = 0.0, tmax = 0.0;
        for (int g = g0; g < g1; ++g) {
            const double T = sU[g - g0];
            tmin = (g == g0) ? T : fmin(tmin, T);
            tmax = (g == g0) ? T : fmax(tmax, T);
        }
        if (fmax(fabs(tmin - tt), fabs(tmax - tt)) < A.fcstol * tt) {
            if (threadIdx.x == 0 && blockIdx.x == A.ub0[s]) A.done[s] = 1;
            return;
        }
    }
    double T3[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (NQ == 4 && kind == 4) {
        for (int q = 0; q < 6; ++q) T3[q] = s4d[6 * prm + q];
    } else if (kind == 3) {
        T3[0] = s3d[6 * prm]; T3[1] = s3d[6 * prm + 1]; T3[2] = s3d[6 * prm + 2];
    } else if (kind == 2) {
        T3[0] = s2d[6 * prm];
    }
    double val[NQ], valn[NQ];
    int gq[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) { val[q] = 0.0; valn[q] = 0.0; gq[q] = -1; }
    const bool shk = A.shk_on[s] != 0;
    if (live) {
        const SysP &P = sp[s];
        const double dt = P.dt;
        double xn[NQ][3], xo[NQ][3], xb[NQ][3], vd[NQ][3];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            if (q >= kind) break;
            const double im = aim[q];
            if (PH == 2) {   // the updated velocity from PH 1 (and k_remotv)
                for (int k = 0; k < 3; ++k) {
                    vd[q][k] = avx[q][k];
                    xn[q][k] = (round == 1 ? ax[q][k] : axo[q][k]) + dt * avx[q][k];
                }
            } else if (round == 1) {   // verlethe.f90:396-404, 509-516
                const double T = shk ? sT[agid[q] - g0] : ts;
                const double itc = 1.0 / T;
                const double chi = sqrt(tt * itc);
                const double rt = chi * dt * JP_FDCVCV, rt2 = 2.0 * chi - 1.0;
                for (int k = 0; k < 3; ++k) {
                    const double vk = rt2 * avx[q][k] - rt * ag[q][k] * im;
                    vd[q][k] = vk;
                    xn[q][k] = ax[q][k] + dt * vk;
                }
            } else {   // verlethe.f90:611-657: constraint force, then the update
                const double itc = 1.0 / sT[agid[q] - g0];
                const double chi = sqrt(tt * itc);
                const double rt = chi * dt, rt2 = 2.0 * chi - 1.0, r = 1.0 / (dt * dt);
                for (int k = 0; k < 3; ++k) {
                    const double f = afc[q][k] + r * (ax[q][k] - avx[q][k]);
                    afc[q][k] = f;
                    A.fcc[3 * a[q] + k] = f;
                    const double vk = (-JP_FDCVCV * ag[q][k] * im + f) * rt + avb[q][k] * rt2;
                    vd[q][k] = vk;
                    xn[q][k] = axo[q][k] + dt * vk;
                }
            }
            if (PH == 1) {
                for (int k = 0; k < 3; ++k) A.v[3 * a[q] + k] = vd[q][k];
                continue;
            }
            for (int k = 0; k < 3; ++k) {
                xb[q][k] = xn[q][k];
                xo[q][k] = axo[q][k];
                if (shk) A.xcb[3 * a[q] + k] = xn[q][k];
            }
        }
        if constexpr (NQ == 4) {
            if (PH != 1 && shk && kind == 4) {
                const int e = shake4_regs(P, aim, T3, tol, maxit, xn, xo);
                if (e) atomicOr(err, e);
            }
        }
        if (PH != 1 && shk) {
            if (kind == 3) {
                const int e = shake3_regs(P, aim, T3, tol, maxit, xn, xo);
                if (e) atomicOr(err, e);
            } else if (kind == 2) {
                shake2_regs(P, 1.0 / am[0], 1.0 / am[1], T3[0], xn, xo);
            }
        }
        const double r = 1.0 / dt, r2 = 1.0 / (dt * dt), r3 = 0.5 * dt;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            if (PH == 1 || q >= kind) break;
            const int b = a[q];
            if (shk) {
                double w[3], wn[3];
                for (int k = 0; k < 3; ++k) {
                    const int qq = 3 * b + k;
                    A.x[qq] = xn[q][k];
                    const double vk = (xn[q][k] - xo[q][k]) * r;
                    A.v[qq] = vk;
                    w[k] = (avb[q][k] + vk) * 0.5;
                    const double f = afc[q][k] + r2 * (xn[q][k] - xb[q][k]);
                    wn[k] = r3 * (-JP_FDCVCV * ag[q][k] * aim[q] + f) + avb[q][k];
                }
                val[q] = cano_mw2(am[q], w[0], w[1], w[2]);
                valn[q] = cano_mw2(am[q], wn[0], wn[1], wn[2]);
                gq[q] = agid[q];
            } else {
                for (int k = 0; k < 3; ++k) {
                    A.x[3 * b + k] = xn[q][k];
                    A.v[3 * b + k] = vd[q][k];
                }
            }
            JAtom *j = ys + asl[q];   // as cano_sorted, from the preloaded slot and image
            j->x = xn[q][0] - P.L[0] * (double)aimg[q][0];
            j->y = xn[q][1] - P.L[1] * (double)aimg[q][1];
            j->z = xn[q][2] - P.L[2] * (double)aimg[q][2];
        }
    }
    if (PH != 1 && shk && round < A.icslop) {
        block_partials(val, gq, g0, g1, pu_out, A.nub);
        block_partials(valn, gq, g0, g1, pn_out, A.nub);
    }
}
