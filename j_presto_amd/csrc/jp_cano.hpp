// Canonical MD (md_run METHOD= CANO, verlethe.f90:79-757 with idfvsl = 3,
// constant lambda, STOPCE= NO, HEATLO= 0).
//
// One step after the forces:
//   begin   v_b = v, x_o = x, f_c = 0, w = v - (dt/2) c g / m  -> m|w|^2
//   (group partial sums)
//   drift   chi = sqrt(T_set / T_g(w)) (1 without SHAKE), v = (2 chi - 1) v
//           - chi dt c g / m, x += dt v
//   icslop rounds, each skipped once the system is done:
//     shake   (rounds after the first: chi from w, v = (-c g / m + f_c)
//             chi dt + v_b (2 chi - 1), x = x_o + dt v), then x_b = x,
//             SHAKE(x; x_o), v = (x - x_o)/dt, w = (v_b + v)/2
//     (group partial sums)
//     check   every group within fcstol T_set, or the last round: done;
//             else f_c += (x - x_b)/dt^2, w = (dt/2)(-c g / m + f_c) + v_b
//     (group partial sums, unless last round)
// The sorted coordinates follow x in the shake kernel (in the drift kernel
// for a system without SHAKE groups).
// T_g = (sum over group g of m|w|^2) 1e7 / (dof_g gascst) (caltmp.f90), with
// one group per system (TEMPCO= NO) or per molecule kind and cluster
// (TEMPCO= YES, caltmp2).  c = 4.184e-4 (md.f90:56), 1/m is md_run's ifxmass.
// Group sums: k_seg_part leaves JP_SEG_NB fixed-order partials per group
// (over the atoms sorted by group); every block of a consumer kernel sums the
// partials of the groups its atoms belong to with the same butterfly, so all
// blocks see the same temperatures and runs are deterministic.  The order
// differs from md_run's serial sum, which moves results in the last bits.
// Expressions keep md_run's operand order.
#pragma once
#include "jp_kernels.hpp"

#define JP_GASCST 8.31451   // PHYCNS.f90:21
#define JP_FDCVCV 4.184e-4  // md.f90:56

#define JP_CANO_MAXG 32   // temperature groups one block of a canonical kernel may touch

struct CanoArgs {
    int natom, nunit, icslop;
    double fcstol;
    const int *atom_sys, *inum, *img;
    const int *gid, *gpos;     // temperature group of each atom; its index in group order
    const int *sg0, *sg1;      // groups of system s: [sg0[s], sg1[s])
    const int *sfirst;         // first atom of system s (writes the system's flag)
    const int *shk_on;         // system s has SHAKE groups
    const double *gdof, *temps;
    const double *mass, *imass, *gslot;
    double *x, *v, *xold, *vback, *fcc, *xcb, *kp;
    const double *part;        // per group: JP_SEG_NB partial sums of m|w|^2
    int *done;                 // per system: the SHAKE/temperature rounds have ended
};

static_assert(JP_SEG_NB == 64, "one partial per lane");

// Temperatures of groups [g0, g1) into sT (caltmp.f90:39-43): the 64
// partials of a group reduced by a butterfly (every lane, and every block,
// gets the same value).  All threads of the block take part.
__device__ __forceinline__ void block_temps(const CanoArgs &A, int g0, int g1, double *sT) {
#pragma clang fp contract(off)
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63, nw = blockDim.x >> 6;
    for (int g = g0 + w; g < g1; g += nw) {
        double p = A.part[(size_t)g * JP_SEG_NB + l];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) p = p + __shfl_xor(p, o, 64);
        if (l == 0) sT[g - g0] = p * 1.0e7 / (A.gdof[g] * JP_GASCST);
    }
    __syncthreads();
}

// Groups of the systems of atoms [a0, a1] (a block's atoms).
__device__ __forceinline__ void block_groups(const CanoArgs &A, int a0, int a1, int &g0, int &g1) {
    g0 = A.sg0[A.atom_sys[a0]];
    g1 = A.sg1[A.atom_sys[a1]];
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
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= A.natom) return;
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
    A.kp[A.gpos[a]] = cano_mw2(A.mass[a], w[0], w[1], w[2]);
    if (a == A.sfirst[s]) A.done[s] = A.shk_on[s] ? 0 : 1;
}

__global__ void __launch_bounds__(256) k_cano_drift(CanoArgs A, const SysP *__restrict__ sp, JAtom *__restrict__ ys) {
#pragma clang fp contract(off)
    __shared__ double sT[JP_CANO_MAXG];
    const int b0 = blockIdx.x * blockDim.x, a = b0 + threadIdx.x;
    int g0, g1;
    block_groups(A, b0, min(b0 + (int)blockDim.x, A.natom) - 1, g0, g1);
    block_temps(A, g0, g1, sT);
    if (a >= A.natom) return;
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
    const int u0 = blockIdx.x * blockDim.x, u = u0 + threadIdx.x;
    int g0 = 0, g1 = 0;
    if (update) {
        // a unit's atoms share its system; units follow the atom order
        const int ul = min(u0 + (int)blockDim.x, A.nunit) - 1;
        block_groups(A, unit[u0].x, unit[ul].x, g0, g1);
        block_temps(A, g0, g1, sT);
    }
    if (u >= A.nunit) return;
    const int4 U = unit[u];
    const int kind = U.w & 3, prm = U.w >> 2;
    const int a[3] = {U.x, U.y, U.z};
    const int s = A.atom_sys[a[0]];
    if (A.done[s]) return;
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
        A.kp[A.gpos[b]] = cano_mw2(A.mass[b], w[0], w[1], w[2]);
        cano_sorted(P, A, b, xn[q], ys);
    }
}

// End of round `round` (1-based): every group of the system within fcstol of
// its target, or the last round, ends the rounds (verlethe.f90:601-611);
// otherwise the constraint force grows and w is the next trial velocity.
__global__ void __launch_bounds__(256) k_cano_check(CanoArgs A, const SysP *__restrict__ sp, int round) {
#pragma clang fp contract(off)
    __shared__ double sT[JP_CANO_MAXG];
    const int b0 = blockIdx.x * blockDim.x, a = b0 + threadIdx.x;
    int g0, g1;
    block_groups(A, b0, min(b0 + (int)blockDim.x, A.natom) - 1, g0, g1);
    block_temps(A, g0, g1, sT);
    if (a >= A.natom) return;
    const int s = A.atom_sys[a];
    if (A.done[s]) return;
    const double ts = A.temps[s];
    double tmin = 0.0, tmax = 0.0;
    for (int g = A.sg0[s]; g < A.sg1[s]; ++g) {
        const double T = sT[g - g0];
        tmin = (g == A.sg0[s]) ? T : fmin(tmin, T);
        tmax = (g == A.sg0[s]) ? T : fmax(tmax, T);
    }
    if (fmax(fabs(tmin - ts), fabs(tmax - ts)) < A.fcstol * ts || round == A.icslop) {
        if (a == A.sfirst[s]) A.done[s] = 1;
        return;
    }
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
    A.kp[A.gpos[a]] = cano_mw2(A.mass[a], w[0], w[1], w[2]);
}
