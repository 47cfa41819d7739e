// Multicanonical MD (md_run METHOD= GE, idfvsl = 4, one energy class):
// every step the forces carry a weight from the potential energy of the
// current coordinates (verlethe.f90:196-233),
//   r = (E - celw) / (ceup - celw),
//   dprobl = alphalw (E <= celw), alphaup (E >= ceup), else
//            c(0,i) + sum_j c(j,i) (r - low_v(i))^j over the first window i
//            with r < high_v(i),
//   weight = (gascst/joucal) 1e-3 tempce dprobl,
// with E = epotcc(1,1), dnergy's sum of the energy slots in slot order
// (dnergy.f90:78).  k_ge_weight forms E from the step's device sums (pair
// energies per cluster, excess energies per atom, bonded energies per term)
// in a fixed order and leaves the weight in device memory, where the CANO
// kernels read it; the step never waits for the host.
#pragma once
#include "jp_cano.hpp"

#define JP_JOUCAL 4.184   // PHYCNS.f90:25

struct GeArgs {
    int nw, nd;               // windows, polynomial degree
    const double *par;        // alphalw, alphaup, celw, ceup, tempce, low[nw], high[nw], c[nw][nd + 1]
    const double *rsum;       // FP64 pair sums [6] (elec, vdw per class), excess sums at [6 * nsys]
    const long long *rfix;    // fixed-point pair sums [6] (FP32 path), else null
    int nsys;
    double hf;                // pair sums count each pair twice (0.5) or once (1)
    double dself;             // ZD self energy of the class
    const double *eterm;      // bonded energy of every term
    const int *eslot;         // its energy slot (bond, angle, torsion, improper, 1-4 vdW, 1-4 elec)
    int ne;
    double *out;              // [0] weight, [1] E
};

// x**j for j >= 1 as md_run's real**integer is evaluated (llvm.powi:
// binary exponentiation, squaring the base, multiplying the set bits of j
// from the lowest up).
__device__ __forceinline__ double ge_powi(double a, int b) {
#pragma clang fp contract(off)
    double r = 1.0;
    for (;;) {
        if (b & 1) r = r * a;
        b /= 2;
        if (b == 0) break;
        a = a * a;
    }
    return r;
}

__global__ void __launch_bounds__(256) k_ge_weight(GeArgs G) {
#pragma clang fp contract(off)
    __shared__ double s_e[6][256];
    const int t = threadIdx.x;
    // bonded energies per slot: strided partial sums, then a fixed tree
    double p[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int q = t; q < G.ne; q += 256) {
        const int k = G.eslot[q];
        const double e = G.eterm[q];
#pragma unroll
        for (int m = 0; m < 6; ++m) p[m] = p[m] + (k == m ? e : 0.0);
    }
#pragma unroll
    for (int m = 0; m < 6; ++m) s_e[m][t] = p[m];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h)
#pragma unroll
            for (int m = 0; m < 6; ++m) s_e[m][t] = s_e[m][t] + s_e[m][t + h];
        __syncthreads();
    }
    if (t != 0) return;
    double el, vd;
    if (G.rfix) {
        el = (double)G.rfix[0] / JP_ESCALE;
        vd = (double)G.rfix[3] / JP_ESCALE;
    } else {
        el = G.rsum[0];
        vd = G.rsum[3];
    }
    el = el * G.hf + G.dself + G.rsum[6 * (size_t)G.nsys];
    vd = vd * G.hf;
    double E = 0.0;
    for (int m = 0; m < 6; ++m) E = E + s_e[m][0];
    E = E + vd;
    E = E + el;
    const double *par = G.par;
    const double alphalw = par[0], alphaup = par[1], celw = par[2], ceup = par[3], tempce = par[4];
    const double *low = par + 5, *high = par + 5 + G.nw, *c = par + 5 + 2 * G.nw;
    double dprobl;
    if (E <= celw) {
        dprobl = alphalw;
    } else if (E >= ceup) {
        dprobl = alphaup;
    } else {
        double r = (E - celw) / (ceup - celw);
        // md_run reads past its window arrays for r >= the last high_v: the
        // last window is taken there
        int i = 0;
        while (i < G.nw - 1 && !(r < high[i])) ++i;
        r = r - low[i];
        const double *ci = c + (size_t)i * (G.nd + 1);
        dprobl = ci[0];
        for (int j = 1; j <= G.nd; ++j) dprobl = dprobl + ci[j] * ge_powi(r, j);
    }
    double dew = (JP_GASCST / JP_JOUCAL) * 1.0e-3;
    dew = dew * tempce * dprobl;
    G.out[0] = dew;
    G.out[1] = E;
}
