// Canonical MD (md_run METHOD= CANO, verlethe.f90:79-757 with idfvsl = 3,
// constant lambda, STOPCE= NO, HEATLO= 0).
//
// One step after the forces:
//   begin   v_b = v, x_o = x, f_c = 0, w = v - (dt/2) c g / m  -> m|w|^2
//   (group sums)
//   drift   chi = sqrt(T_set / T_g(w)) (1 without SHAKE), v = (2 chi - 1) v
//           - chi dt c g / m, x += dt v
//   icslop rounds, each skipped once the system is done:
//     shake   x_b = x, SHAKE(x; x_o), v = (x - x_o)/dt, w = (v_b + v)/2
//     (group sums)
//     check   every group within fcstol T_set, or the last round: done;
//             else f_c += (x - x_b)/dt^2, w = (dt/2)(-c g / m + f_c) + v_b
//     (group sums)
//     update  chi from w, v = (-c g / m + f_c) chi dt + v_b (2 chi - 1),
//             x = x_o + dt v
//   end     sorted coordinates from x
// T_g = (sum over group g of m|w|^2) 1e7 / (dof_g gascst) (caltmp.f90), with
// one group per system (TEMPCO= NO) or per molecule kind and cluster
// (TEMPCO= YES, caltmp2).  c = 4.184e-4 (md.f90:56), 1/m is md_run's ifxmass.
// Group sums are fixed-order (k_seg_part / k_seg_fin over the atoms sorted by
// group), so runs are deterministic; their order differs from md_run's serial
// sum, which moves results in the last bits.  Expressions keep md_run's
// operand order.
#pragma once
#include "jp_kernels.hpp"

#define JP_GASCST 8.31451   // PHYCNS.f90:21
#define JP_FDCVCV 4.184e-4  // md.f90:56

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
    const double *gk;          // per group: sum of m|w|^2
    int *done;                 // per system: the SHAKE/temperature rounds have ended
};

__device__ __forceinline__ double cano_temp(const CanoArgs &A, int g) {
#pragma clang fp contract(off)
    return A.gk[g] * 1.0e7 / (A.gdof[g] * JP_GASCST);
}

__device__ __forceinline__ double cano_mw2(double m, double w0, double w1, double w2) {
#pragma clang fp contract(off)
    return m * (w0 * w0 + w1 * w1 + w2 * w2);
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

__global__ void __launch_bounds__(256) k_cano_drift(CanoArgs A, const SysP *__restrict__ sp) {
#pragma clang fp contract(off)
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= A.natom) return;
    const int s = A.atom_sys[a];
    const double dt = sp[s].dt, im = A.imass[a], ts = A.temps[s];
    const double T = A.shk_on[s] ? cano_temp(A, A.gid[a]) : ts;   // verlethe.f90:396-404
    const double itc = 1.0 / T;
    const double chi = sqrt(ts * itc);
    const double rt = chi * dt * JP_FDCVCV, rt2 = 2.0 * chi - 1.0;
    const size_t sl = (size_t)A.inum[a];
    for (int k = 0; k < 3; ++k) {
        const double vk = rt2 * A.v[3 * a + k] - rt * A.gslot[3 * sl + k] * im;
        A.v[3 * a + k] = vk;
        A.x[3 * a + k] = A.x[3 * a + k] + dt * vk;
    }
}

// One SHAKE round per unit (SHAKE group or free atom), as k_step's.
__global__ void __launch_bounds__(128) k_cano_shake(CanoArgs A, const SysP *__restrict__ sp,
                                                    const int4 *__restrict__ unit, const double *__restrict__ s2d,
                                                    const double *__restrict__ s3d, double tol, int maxit,
                                                    int *__restrict__ err) {
#pragma clang fp contract(off)
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= A.nunit) return;
    const int4 U = unit[u];
    const int kind = U.w & 3, prm = U.w >> 2;
    const int a[3] = {U.x, U.y, U.z};
    const int s = A.atom_sys[a[0]];
    if (A.done[s]) return;
    const SysP &P = sp[s];
    double xn[3][3], xo[3][3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        if (q >= kind) break;
        for (int k = 0; k < 3; ++k) {
            const double p = A.x[3 * a[q] + k];
            A.xcb[3 * a[q] + k] = p;
            xn[q][k] = p;
            xo[q][k] = A.xold[3 * a[q] + k];
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
    const double r = 1.0 / P.dt;
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
    }
}

// End of round `round` (1-based): every group of the system within fcstol of
// its target, or the last round, ends the rounds (verlethe.f90:601-611);
// otherwise the constraint force grows and w is the next trial velocity.
__global__ void __launch_bounds__(256) k_cano_check(CanoArgs A, const SysP *__restrict__ sp, int round) {
#pragma clang fp contract(off)
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= A.natom) return;
    const int s = A.atom_sys[a];
    if (A.done[s]) return;
    const double ts = A.temps[s];
    double tmin = 0.0, tmax = 0.0;
    for (int g = A.sg0[s]; g < A.sg1[s]; ++g) {
        const double T = cano_temp(A, g);
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

__global__ void __launch_bounds__(256) k_cano_update(CanoArgs A, const SysP *__restrict__ sp) {
#pragma clang fp contract(off)
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= A.natom) return;
    const int s = A.atom_sys[a];
    if (A.done[s]) return;
    const double dt = sp[s].dt, im = A.imass[a];
    const double itc = 1.0 / cano_temp(A, A.gid[a]);
    const double chi = sqrt(A.temps[s] * itc);
    const double rt = chi * dt, rt2 = 2.0 * chi - 1.0;
    const size_t sl = (size_t)A.inum[a];
    for (int k = 0; k < 3; ++k) {
        const int q = 3 * a + k;
        const double vk = (-JP_FDCVCV * A.gslot[3 * sl + k] * im + A.fcc[q]) * rt + A.vback[q] * rt2;
        A.v[q] = vk;
        A.x[q] = A.xold[q] + dt * vk;
    }
}

// Sorted coordinates (the pair kernels' j atoms) from x, as k_step does.
__global__ void __launch_bounds__(256) k_cano_end(CanoArgs A, const SysP *__restrict__ sp, JAtom *__restrict__ ys) {
#pragma clang fp contract(off)
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= A.natom) return;
    const SysP &P = sp[A.atom_sys[a]];
    JAtom *j = ys + A.inum[a];
    j->x = A.x[3 * a] - P.L[0] * (double)A.img[3 * a];
    j->y = A.x[3 * a + 1] - P.L[1] * (double)A.img[3 * a + 1];
    j->z = A.x[3 * a + 2] - P.L[2] * (double)A.img[3 * a + 2];
}
