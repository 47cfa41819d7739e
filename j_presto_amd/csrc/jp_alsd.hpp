// Adaptive lambda square dynamics (md_run METHOD= GE under GEPSME= LMD:
// idfvsl = 4 with cluster_method LMD, verlethe.f90:235-290).  lambda, the
// scaling of the class-1 (lambda^2) and class-2 (lambda) energies, is a
// dynamical variable of mass m_l that moves beside the atoms under
//   lambda_f = -2 E1 lambda - E2 - dew(lambda),
// E1, E2 the unscaled class-1 and class-2 potential energies of the step
// (dnergy's slot sums, dnergy.f90:78) and dew the bias of the NAMEGE=
// windows taken in lambda itself (the first window whose high_v exceeds
// lambda, else the last; its polynomial about low_v) plus the walls
// lower * (celw - lambda) below celw and upper * (lambda - ceup) above ceup,
// times (gascst/joucal) 1e-3 tempce.
//
// Device state per system (JP_LS doubles), so that a step never waits for
// the host and replays as a graph:
//   [0] lambda, [1] its velocity, [2] lambda_f of the step, [3] the
//   half-step velocity lambda_v + (dt/2) c lambda_f / m_l (round-1
//   temperatures), [4] lambda and [5] its velocity at the step start,
//   [6] (lambda_v at the start + after the drift) / 2 (the temperatures of
//   later SHAKE rounds), [7] E1, [8] E2, [9] dew.
// k_alsd_fac turns lambda into the class factors (lambda^2, lambda, 1) the
// pair, excess and bonded kernels read; k_alsd forms E1, E2 and lambda_f
// after the forces; the canonical kernels carry lambda's kinetic term in
// the temperatures and move lambda in round 1 (jp_cano.hpp).
#pragma once
#include "jp_ge.hpp"

// lambda -> the class factors of every cluster pair (SysP::fac) and of the
// excess and bonded terms (efac), as the host's factors under LMD.
__global__ void k_alsd_fac(int nsys, const double *__restrict__ lam, SysP *__restrict__ sp,
                           double *__restrict__ efac) {
#pragma clang fp contract(off)
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    const double l = lam[JP_LS * s];
    const double f[3] = {l * l, l, 1.0};
    SysP &P = sp[s];
#pragma unroll
    for (int q = 0; q < 4; ++q) P.fac[q] = f[P.cls[q]];
#pragma unroll
    for (int k = 0; k < 3; ++k) efac[3 * s + k] = f[k];
}

struct AlsdArgs {
    int nw, nd;               // windows, polynomial degree
    const double *par;        // lower, upper, celw, ceup, tempce, low[nw], high[nw], c[nw][nd + 1]
    const double *rsum;       // FP64 pair sums [s][6] (elec, vdw per class), excess sums [s][3] at [6 nsys]
    const long long *rfix;    // fixed-point pair sums [s][6] (FP32 path), else null
    int nsys;
    double hf;                // pair sums count each pair twice (0.5) or once (1)
    const double *dself;      // [s][3] ZD self energy per class
    const double *eterm;      // bonded energy of every term
    const int *esys, *ecls, *eslot;   // its system, class and slot (bond ... 1-4 elec)
    int ne;
    const double *lmass;      // [s] lambda's mass
    double *lam;              // [s][JP_LS] state
    const SysP *sp;
};

// One block per system: the class-1 and class-2 energies (bonded terms per
// slot in strided partial sums and a fixed tree, then VDW15 and ELE15 as
// dnergy orders the slots), the bias and lambda's force; the step-start
// state and the half-step velocity.
__global__ void __launch_bounds__(256) k_alsd(AlsdArgs A) {
#pragma clang fp contract(off)
    __shared__ double s_e[12][256];
    const int t = threadIdx.x, s = blockIdx.x;
    double p[12];
#pragma unroll
    for (int m = 0; m < 12; ++m) p[m] = 0.0;
    for (int q = t; q < A.ne; q += 256) {
        if (A.esys[q] != s) continue;
        const int c = A.ecls[q];
        if (c > 1) continue;
        const int k = 6 * c + A.eslot[q];
        const double e = A.eterm[q];
#pragma unroll
        for (int m = 0; m < 12; ++m) p[m] = p[m] + (k == m ? e : 0.0);
    }
#pragma unroll
    for (int m = 0; m < 12; ++m) s_e[m][t] = p[m];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h)
#pragma unroll
            for (int m = 0; m < 12; ++m) s_e[m][t] = s_e[m][t] + s_e[m][t + h];
        __syncthreads();
    }
    if (t != 0) return;
    double E[2];
    for (int c = 0; c < 2; ++c) {
        double el, vd;
        if (A.rfix) {
            el = (double)A.rfix[6 * s + c] / JP_ESCALE;
            vd = (double)A.rfix[6 * s + 3 + c] / JP_ESCALE;
        } else {
            el = A.rsum[6 * s + c];
            vd = A.rsum[6 * s + 3 + c];
        }
        el = el * A.hf + A.dself[3 * s + c] + A.rsum[6 * (size_t)A.nsys + 3 * s + c];
        vd = vd * A.hf;
        double e = 0.0;
        for (int m = 0; m < 6; ++m) e = e + s_e[6 * c + m][0];
        e = e + vd;
        e = e + el;
        E[c] = e;
    }
    double *L = A.lam + JP_LS * s;
    const double lambda = L[0], lv = L[1];
    const double *par = A.par;
    const double lower = par[0], upper = par[1], celw = par[2], ceup = par[3], tempce = par[4];
    const double *low = par + 5, *high = par + 5 + A.nw, *c = par + 5 + 2 * A.nw;
    int i = A.nw - 1;
    for (int j = 0; j < A.nw; ++j)
        if (lambda < high[j]) { i = j; break; }
    const double rtmp = lambda - low[i];
    const double *ci = c + (size_t)i * (A.nd + 1);
    double dprobl = ci[0];
    for (int j = 1; j <= A.nd; ++j) dprobl = dprobl + ci[j] * ge_powi(rtmp, j);
    if (lambda < celw) dprobl = dprobl + lower * (celw - lambda);
    else if (lambda > ceup) dprobl = dprobl + upper * (lambda - ceup);
    const double dew = (JP_GASCST / JP_JOUCAL) * 1.0e-3 * tempce * dprobl;
    const double lf = -(2.0 * E[0] * lambda) - E[1] - dew;
    const double r = 0.5 * A.sp[s].dt * JP_FDCVCV;
    L[2] = lf;
    L[3] = lv + r * lf * (1.0 / A.lmass[s]);
    L[4] = lambda;
    L[5] = lv;
    L[7] = E[0];
    L[8] = E[1];
    L[9] = dew;
}
