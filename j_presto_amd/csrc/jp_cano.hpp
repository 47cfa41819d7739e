// Canonical MD (md_run METHOD= CANO, verlethe.f90:79-757 with idfvsl = 3,
// constant lambda, STOPCE= NO, HEATLO= 0).
//
// One step after the forces:
//   begin   v_b = v, x_o = x, f_c = 0, w = v - (dt/2) c g / m  -> m|w|^2
//   drift   chi = sqrt(T_set / T_g(w)) (1 without SHAKE), v = (2 chi - 1) v
//           - chi dt c g / m, x += dt v
//   icslop rounds, each skipped once the system is done:
//     shake   (rounds after the first: chi from w, v = (-c g / m + f_c)
//             chi dt + v_b (2 chi - 1), x = x_o + dt v), then x_b = x,
//             SHAKE(x; x_o), v = (x - x_o)/dt, w = (v_b + v)/2
//     check   every group within fcstol T_set, or the last round: done;
//             else f_c += (x - x_b)/dt^2, w = (dt/2)(-c g / m + f_c) + v_b
// The sorted coordinates follow x in the shake kernel (in the drift kernel
// for a system without SHAKE groups).
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

struct CanoArgs {
    int natom, nunit, icslop, nab, nub;   // nab, nub: atom blocks, unit blocks
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
    const double *mass, *imass, *gslot;
    double *x, *v, *xold, *vback, *fcc, *xcb;
    double *pa, *pu;           // partials of m|w|^2: [group][atom block], [group][unit block]
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

// Temperatures of groups [g0, g1) into sT (caltmp.f90:39-43) from the
// partials of the blocks [bl0[s], bl1[s]) of each group's system: lane l adds
// blocks l, l+64, ... in order, then a butterfly (every block gets the same
// value).  All threads of the block take part.
__device__ __forceinline__ void block_temps(const CanoArgs &A, const double *__restrict__ P, int nblk,
                                            const int *__restrict__ bl0, const int *__restrict__ bl1, int g0, int g1,
                                            double *sT) {
#pragma clang fp contract(off)
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63, nw = blockDim.x >> 6;
    for (int g = g0 + w; g < g1; g += nw) {
        const int s = A.gsys[g];
        double p = 0.0;
        for (int b = bl0[s] + l; b < bl1[s]; b += 64) p = p + P[(size_t)g * nblk + b];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) p = p + __shfl_xor(p, o, 64);
        if (l == 0) sT[g - g0] = p * 1.0e7 / (A.gdof[g] * JP_GASCST);
    }
    __syncthreads();
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
            w[k] = vk - r * A.gslot[3 * sl + k] * im;
        }
        val[0] = cano_mw2(A.mass[a], w[0], w[1], w[2]);
        gq[0] = A.gid[a];
        if (a == A.sfirst[s]) A.done[s] = A.shk_on[s] ? 0 : 1;
    }
    int g0, g1;
    block_groups(A, B.x, g0, g1);
    block_partials(val, gq, g0, g1, A.pa, A.nab);
}

__global__ void __launch_bounds__(256) k_cano_drift(CanoArgs A, const SysP *__restrict__ sp, JAtom *__restrict__ ys) {
#pragma clang fp contract(off)
    __shared__ double sT[JP_CANO_MAXG];
    const int2 B = A.ablk[blockIdx.x];
    const int a = B.x + threadIdx.x;
    int g0, g1;
    block_groups(A, B.x, g0, g1);
    block_temps(A, A.pa, A.nab, A.ab0, A.ab1, g0, g1, sT);
    if (a >= B.y) return;
    const int s = A.atom_sys[a];
    const SysP &P = sp[s];
    const double dt = P.dt, im = A.imass[a], ts = A.temps[s];
    const double T = A.shk_on[s] ? sT[A.gid[a] - g0] : ts;   // verlethe.f90:396-404
    const double itc = 1.0 / T;
    const double chi = sqrt(ts * itc);
    const double rt = chi * dt * JP_FDCVCV, rt2 = 2.0 * chi - 1.0;
    const size_t sl = (size_t)A.inum[a];
    double xa[3];
    for (int k = 0; k < 3; ++k) {
        const double vk = rt2 * A.v[3 * a + k] - rt * A.gslot[3 * sl + k] * im;
        A.v[3 * a + k] = vk;
        xa[k] = A.x[3 * a + k] + dt * vk;
        A.x[3 * a + k] = xa[k];
    }
    if (!A.shk_on[s]) cano_sorted(P, A, a, xa, ys);
}

// One SHAKE round per unit (SHAKE group or free atom), as k_step's; rounds
// after the first first apply the previous round's velocity update.
__global__ void __launch_bounds__(128) k_cano_shake(CanoArgs A, const SysP *__restrict__ sp,
                                                    const int4 *__restrict__ unit, const double *__restrict__ s2d,
                                                    const double *__restrict__ s3d, double tol, int maxit,
                                                    int update, JAtom *__restrict__ ys, int *__restrict__ err) {
#pragma clang fp contract(off)
    __shared__ double sT[JP_CANO_MAXG];
    const int2 B = A.ublk[blockIdx.x];
    const int u = B.x + threadIdx.x;
    // a unit's atoms share its system
    int g0, g1;
    block_groups(A, unit[B.x].x, g0, g1);
    if (update) block_temps(A, A.pa, A.nab, A.ab0, A.ab1, g0, g1, sT);
    double val[3] = {0.0, 0.0, 0.0};
    int gq[3] = {-1, -1, -1};
    const int4 U = unit[min(u, B.y - 1)];
    const int kind = U.w & 3, prm = U.w >> 2;
    const int a[3] = {U.x, U.y, U.z};
    const int s = A.atom_sys[a[0]];
    if (u < B.y && !A.done[s]) {
        const SysP &P = sp[s];
        const double dt = P.dt;
        double xn[3][3], xo[3][3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            if (q >= kind) break;
            const int b = a[q];
            if (update) {   // verlethe.f90:646-657 of the previous round
                const double itc = 1.0 / sT[A.gid[b] - g0];
                const double chi = sqrt(A.temps[s] * itc);
                const double rt = chi * dt, rt2 = 2.0 * chi - 1.0, im = A.imass[b];
                const size_t sl = (size_t)A.inum[b];
                for (int k = 0; k < 3; ++k) {
                    const int qq = 3 * b + k;
                    const double vk = (-JP_FDCVCV * A.gslot[3 * sl + k] * im + A.fcc[qq]) * rt + A.vback[qq] * rt2;
                    A.v[qq] = vk;
                    A.x[qq] = A.xold[qq] + dt * vk;
                }
            }
            for (int k = 0; k < 3; ++k) {
                const double p = A.x[3 * b + k];
                A.xcb[3 * b + k] = p;
                xn[q][k] = p;
                xo[q][k] = A.xold[3 * b + k];
            }
        }
        if (kind == 3) {
            const double im[3] = {A.imass[a[0]], A.imass[a[1]], A.imass[a[2]]};
            const double T[3] = {s3d[6 * prm], s3d[6 * prm + 1], s3d[6 * prm + 2]};
            const int e = shake3_regs(P, im, T, tol, maxit, xn, xo);
            if (e) atomicOr(err, e);
        } else if (kind == 2) {
            shake2_regs(P, 1.0 / A.mass[a[0]], 1.0 / A.mass[a[1]], s2d[6 * prm], xn, xo);
        }
        const double r = 1.0 / dt;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            if (q >= kind) break;
            const int b = a[q];
            double w[3];
            for (int k = 0; k < 3; ++k) {
                A.x[3 * b + k] = xn[q][k];
                const double vk = (xn[q][k] - xo[q][k]) * r;
                A.v[3 * b + k] = vk;
                w[k] = (A.vback[3 * b + k] + vk) * 0.5;
            }
            val[q] = cano_mw2(A.mass[b], w[0], w[1], w[2]);
            gq[q] = A.gid[b];
            cano_sorted(P, A, b, xn[q], ys);
        }
    }
    block_partials(val, gq, g0, g1, A.pu, A.nub);
}

// End of round `round` (1-based): every group of the system within fcstol of
// its target, or the last round, ends the rounds (verlethe.f90:601-611);
// otherwise the constraint force grows and w is the next trial velocity.
__global__ void __launch_bounds__(256) k_cano_check(CanoArgs A, const SysP *__restrict__ sp, int round) {
#pragma clang fp contract(off)
    __shared__ double sT[JP_CANO_MAXG];
    const int2 B = A.ablk[blockIdx.x];
    const int a = B.x + threadIdx.x;
    int g0, g1;
    block_groups(A, B.x, g0, g1);
    block_temps(A, A.pu, A.nub, A.ub0, A.ub1, g0, g1, sT);
    double val[1] = {0.0};
    int gq[1] = {-1};
    const int s = A.atom_sys[B.x];
    if (a < B.y && !A.done[s]) {
        const double ts = A.temps[s];
        double tmin = 0.0, tmax = 0.0;
        for (int g = A.sg0[s]; g < A.sg1[s]; ++g) {
            const double T = sT[g - g0];
            tmin = (g == A.sg0[s]) ? T : fmin(tmin, T);
            tmax = (g == A.sg0[s]) ? T : fmax(tmax, T);
        }
        if (fmax(fabs(tmin - ts), fabs(tmax - ts)) < A.fcstol * ts || round == A.icslop) {
            if (a == A.sfirst[s]) A.done[s] = 1;
        } else {
            const double dt = sp[s].dt, im = A.imass[a];
            const double r = 1.0 / (dt * dt), r2 = 0.5 * dt;
            const size_t sl = (size_t)A.inum[a];
            double w[3];
            for (int k = 0; k < 3; ++k) {
                const int q = 3 * a + k;
                const double f = A.fcc[q] + r * (A.x[q] - A.xcb[q]);
                A.fcc[q] = f;
                w[k] = r2 * (-JP_FDCVCV * A.gslot[3 * sl + k] * im + f) + A.vback[q];
            }
            val[0] = cano_mw2(A.mass[a], w[0], w[1], w[2]);
            gq[0] = A.gid[a];
        }
    }
    if (round < A.icslop) block_partials(val, gq, g0, g1, A.pa, A.nab);
}
