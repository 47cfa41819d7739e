// Centre-of-mass stopping (md_run STOPCE, remotv.f90:2-131) and the lone
// kick of the steps that stop it.
//
// A step with STOPCE takes the motion of the first chains out of the
// velocities between the kick and the drift: MICR as k_kick -> k_remotv ->
// k_step<false, true> (verlet.f90:144-190), the canonical step between the
// velocity update and the drift of every round (verlethe.f90:489-502,
// 705-718), and md_run's initial step of a run without a restart as
// k_kick(half) -> k_remotv -> k_step<false, false> (staver.f90:158-207).
#pragma once
#include "jp_kernels.hpp"

#define JP_STOP_NT 256
#define JP_ERR_STOP_MASS 16   // remotv.f90:47-50: total mass <= 0
#define JP_ERR_STOP_SING 32   // remotv.f90:86-88: singular inertia tensor

// The kick alone, per atom: verlet.f90:144-153, v + (-dt c) g / m (vback,
// when given, keeps v before it), or with half staver.f90:158-162,
// v + (-dt c / 2) g (1/m).  g is the step's gradient slot.
__global__ void __launch_bounds__(256) k_kick(int natom, const int *__restrict__ atom_sys,
                                              const SysP *__restrict__ sp, const int *__restrict__ inum,
                                              const double *__restrict__ gslot, const double *__restrict__ mass,
                                              const double *__restrict__ imass, int half, double *__restrict__ v,
                                              double *__restrict__ vback) {
#pragma clang fp contract(off)
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natom) return;
    const SysP &P = sp[atom_sys[a]];
    const size_t sl = (size_t)inum[a];
    if (half) {
        const double rt = -0.5 * P.dt * 4.184e-4, im = imass[a];   // md.f90:56 fdcvcv
        for (int k = 0; k < 3; ++k) v[3 * a + k] = v[3 * a + k] + rt * gslot[3 * sl + k] * im;
    } else {
        const double rt = -P.dt * 4.184e-4, m = mass[a];
        for (int k = 0; k < 3; ++k) {
            const double vk = v[3 * a + k];
            if (vback) vback[3 * a + k] = vk;
            v[3 * a + k] = vk + rt * gslot[3 * sl + k] / m;
        }
    }
}

// Fixed-order sum over the block of NV values per thread: the threads'
// values in LDS, a halving tree; every thread gets the totals in out.
template <int NV>
__device__ __forceinline__ void stop_block_sum(double (*red)[JP_STOP_NT], const double *val, double *out) {
#pragma clang fp contract(off)
    const int t = threadIdx.x;
    for (int k = 0; k < NV; ++k) red[k][t] = val[k];
    __syncthreads();
    for (int h = JP_STOP_NT / 2; h > 0; h >>= 1) {
        if (t < h)
            for (int k = 0; k < NV; ++k) red[k][t] = red[k][t] + red[k][t + h];
        __syncthreads();
    }
    for (int k = 0; k < NV; ++k) out[k] = red[k][0];
    __syncthreads();
}

// remotv for system blockIdx.x: stop = {mode (1 TRAN, 2 ROTA, 3 BOTH; 0
// off), first atom, atoms stopped, 0}.  xp holds x(t) (md_run's cordm1, or
// cord before the drift: the same values).  The centre of mass, the
// momentum and, for ROTA and BOTH, the inertia tensor about the centre of
// mass and the angular momentum as fixed-order block sums (thread t adds
// atoms t, t + 256, ... in order, then a tree), the inverse by cofactors
// (remotv.f90:86-103), then per atom v - w x d - v_cm in md_run's order
// (remotv.f90:120-127).  md_run sums m (invmat (d x v)) atom by atom; here
// invmat multiplies the summed m (d x v), the same up to rounding.  With
// done (the canonical rounds) a round past the first skips systems whose
// rounds have ended.  Errors: JP_ERR_STOP_MASS, JP_ERR_STOP_SING in err, v
// untouched.
__global__ void __launch_bounds__(JP_STOP_NT) k_remotv(const int4 *__restrict__ stop,
                                                       const double *__restrict__ mass,
                                                       const double *__restrict__ xp, double *__restrict__ v,
                                                       const int *__restrict__ done, int round,
                                                       int *__restrict__ err) {
#pragma clang fp contract(off)
    __shared__ double red[9][JP_STOP_NT];
    const int4 S = stop[blockIdx.x];
    const int mode = S.x, a0 = S.y, n = S.z, t = threadIdx.x;
    if (mode == 0 || (done && round > 1 && done[blockIdx.x])) return;   // uniform over the block
    double p[9], tot[9];
    for (int k = 0; k < 7; ++k) p[k] = 0.0;
    for (int i = t; i < n; i += JP_STOP_NT) {
        const int a = a0 + i;
        const double m = mass[a];
        p[0] = p[0] + m;
        for (int k = 0; k < 3; ++k) {
            p[1 + k] = p[1 + k] + m * xp[3 * a + k];
            p[4 + k] = p[4 + k] + m * v[3 * a + k];
        }
    }
    stop_block_sum<7>(red, p, tot);
    if (!(tot[0] > 0.0)) {
        if (t == 0) atomicOr(err, JP_ERR_STOP_MASS);
        return;
    }
    const double itot = 1.0 / tot[0];
    const bool tran = mode == 1 || mode == 3;
    double cen[3], tv[3], av[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < 3; ++k) {
        cen[k] = tot[1 + k] * itot;
        tv[k] = tran ? tot[4 + k] * itot : 0.0;
    }
    if (mode >= 2) {
        for (int k = 0; k < 9; ++k) p[k] = 0.0;
        for (int i = t; i < n; i += JP_STOP_NT) {
            const int a = a0 + i;
            const double m = mass[a];
            const double dx = xp[3 * a] - cen[0], dy = xp[3 * a + 1] - cen[1], dz = xp[3 * a + 2] - cen[2];
            const double vx = v[3 * a], vy = v[3 * a + 1], vz = v[3 * a + 2];
            p[0] = p[0] + m * (dy * dy + dz * dz);
            p[1] = p[1] + m * (dz * dz + dx * dx);
            p[2] = p[2] + m * (dx * dx + dy * dy);
            p[3] = p[3] - m * dx * dy;
            p[4] = p[4] - m * dy * dz;
            p[5] = p[5] - m * dz * dx;
            p[6] = p[6] + m * (dy * vz - dz * vy);
            p[7] = p[7] + m * (dz * vx - dx * vz);
            p[8] = p[8] + m * (dx * vy - dy * vx);
        }
        stop_block_sum<9>(red, p, tot);
        const double ixx = tot[0], iyy = tot[1], izz = tot[2], ixy = tot[3], iyz = tot[4], izx = tot[5];
        const double ixy2 = ixy * ixy, iyz2 = iyz * iyz, izx2 = izx * izx;
        double det = ixx * iyy * izz + 2.0 * ixy * iyz * izx - iyz2 * ixx - izx2 * iyy - ixy2 * izz;
        if (det == 0.0) {
            if (t == 0) atomicOr(err, JP_ERR_STOP_SING);
            return;
        }
        det = 1.0 / det;
        const double i11 = (iyy * izz - iyz2) * det, i12 = (iyz * izx - izz * ixy) * det;
        const double i13 = (ixy * iyz - izx * iyy) * det, i22 = (izz * ixx - izx2) * det;
        const double i23 = (izx * ixy - ixx * iyz) * det, i33 = (ixx * iyy - ixy2) * det;
        av[0] = i11 * tot[6] + i12 * tot[7] + i13 * tot[8];
        av[1] = i12 * tot[6] + i22 * tot[7] + i23 * tot[8];
        av[2] = i13 * tot[6] + i23 * tot[7] + i33 * tot[8];
    }
    for (int i = t; i < n; i += JP_STOP_NT) {
        const int a = a0 + i;
        const double dx = xp[3 * a] - cen[0], dy = xp[3 * a + 1] - cen[1], dz = xp[3 * a + 2] - cen[2];
        v[3 * a] = v[3 * a] - av[1] * dz + av[2] * dy - tv[0];
        v[3 * a + 1] = v[3 * a + 1] - av[2] * dx + av[0] * dz - tv[1];
        v[3 * a + 2] = v[3 * a + 2] - av[0] * dy + av[1] * dx - tv[2];
    }
}
