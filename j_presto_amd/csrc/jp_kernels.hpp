// Device code of libjpgpu: MI355X (gfx950) kernels for j_presto md_run's
// ZD + LJ nonbonded path, the pair-list build that feeds it, and the FP64
// SHAKE / leap-frog step.
//
// Data layout in HBM (one or many systems concatenated; a single system is a
// batch of one):
//   original order  x, v, xold: double[N][3]; per-atom topology arrays
//   sorted "slots"  atoms ordered by (system, box-aligned cell in 3-D Hilbert
//                   order, sub-cell, atom index); every system's slot range
//                   is padded to a multiple of 64; 8 consecutive slots form
//                   one i-cluster.
//     ys[slot]      JAtom {y (FP64, coordinates wrapped at the last rebuild),
//                   q (FP32 chgmod), packed type/class/vdW flags} -- 32 B
//     qd[slot]      chgmod in FP64 (plain FP64 path)
//     xr[slot]      raw coordinates at the last rebuild (list test)
//     lp[slot]      list-test record: FP32 coordinates about the box centre,
//                   bits of w: molecule id (0-29), has-vdW (30); -1 = padding
//   per i-cluster   box, centre, and a j-list of uint2 entries {j slot | image
//                   shift, 8-bit mask} (jp_list.hpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define JP_WAVE 64
#define JP_FSCALE 1099511627776.0   /* 2^40: force fixed point, +-8.4e6 kcal/mol/A */
#define JP_ESCALE 68719476736.0     /* 2^36: energy fixed point, +-1.3e8 kcal/mol */
#define JP_SLOT_BITS 27
#define JP_SLOT_MASK ((1u << JP_SLOT_BITS) - 1u)

struct SysP {
    double L[3], invL[3], wall[3];   // fxcell, invcel, celwal lower walls
    double csz[3], icsz[3];          // box-aligned cell edge and its inverse
    double rorg[3], rr;              // md_run's cell grid: origin, 1/siz (celpar.f90:93-122)
    double rlim2, rc2;               // list radius^2, cutoff^2
    double irlim2;                   // 1/rc^2 (md_run names it irlim2)
    double zcore, bcoeff, fcoeff;    // ZD constants (incel.f90:655-658)
    double fac[4];                   // pair force/energy factor, index ci*2+cj (cluster classes)
    int cls[4];                      // energy class of (ci,cj) (n_matrix - 1)
    int nc[3];
    int cell0, ncell;                // global cell range
    int atom0, natom;                // original atom range
    int slot0, nslot;                // padded slot range
    int lj0, ntyp;                   // LJ table offset / size
    int hbits;                       // nc = 2^hbits cells per edge
    int small;                       // some edge < 2*rlim: minimum image per pair
    int nblim;                       // n15mx: partners of one table allowed per atom (cutoff.f90:30-46)
    int mnv, ipmax, mc0;             // md_run's cells per edge, atoms per cell and class (celpar.f90:182-192),
                                     // offset of this system's per-cell counters
    double dt;
    double ctr[3];                   // box centre: origin of the FP32 list-test frame
    double band;                     // A^2: FP32 list-test results within band of rlim^2 take the FP64 test
    int ctab;                        // bit a: md_run's neighbour-cell table drops pairs within rlim on axis a
    int ctab_sym;                    // 1: those drops are symmetric in i and j (a half list can hold them)
};

struct JAtom {          // 32 B
    double x, y, z;
    float q;
    int t;              // bits 0-11 type, bit 12 has-vdW, bit 13 cluster class
};


__device__ __forceinline__ double anint_d(double x) {
    // Fortran ANINT (round half away from zero); x - trunc(x) is exact
    double t = trunc(x);
    double r = x - t;
    return (fabs(r) >= 0.5) ? t + copysign(1.0, x) : t;
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & (JP_WAVE - 1); }

// XCD-aware block -> work-item mapping: the 8 XCDs take blocks round-robin;
// give each XCD a contiguous range so neighbouring clusters share its L2.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
    int per = (nb + 7) >> 3;
    int xcd = b & 7, k = b >> 3;
    int r = xcd * per + k;
    return r;  // may be >= nb: caller returns
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, JP_WAVE);
    return v;
}

__device__ __forceinline__ long long f2fix(double v, double scale) { return __double2ll_rn(v * scale); }

// cnt[idx] += 1 for every active lane, with one atomic per distinct idx in
// the wavefront (neighbouring atoms share cells, and same-address atomics
// serialize): lanes are grouped by idx first, then every group's first lane
// adds the group's size, all in one atomic instruction.  Returns the lane's
// old value: its group's base plus its rank among the group's lanes.
__device__ __forceinline__ int wave_inc(int *__restrict__ cnt, int idx, bool act) {
    const int l = (int)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
    uint64_t todo = __ballot(act);
    int lead_of = l, rank = 0, size = 0;
    while (todo) {
        const int lead = (int)__builtin_ctzll(todo);
        const int key = __builtin_amdgcn_readlane(idx, lead);
        const uint64_t grp = __ballot(act && idx == key) & todo;
        if ((grp >> l) & 1ull) {
            lead_of = lead;
            rank = (int)__popcll(grp & ((1ull << l) - 1ull));
        }
        if (l == lead) size = (int)__popcll(grp);
        todo &= ~grp;
    }
    int base = 0;
    if (act && lead_of == l) base = atomicAdd(cnt + idx, size);
    base = __shfl(base, lead_of, 64);
    return base + rank;
}

// Index of cell (x,y,z) along the 3-D Hilbert curve of a 2^bits grid
// (J. Skilling, "Programming the Hilbert curve", AIP Conf. Proc. 707, 2004:
// axes -> transpose, then bit interleave).  Consecutive indices are face
// neighbours and the curve is nested: index(2c+s) >> 3 == index(c) one level
// up, so a cell's 2x2x2 sub-cells are visited together.
__host__ __device__ inline uint32_t hilbert3(uint32_t x, uint32_t y, uint32_t z, int bits) {
    uint32_t X[3] = {x, y, z};
    const uint32_t M = 1u << (bits - 1);
    for (uint32_t Q = M; Q > 1; Q >>= 1) {
        const uint32_t P = Q - 1;
        for (int i = 0; i < 3; ++i) {
            if (X[i] & Q) {
                X[0] ^= P;
            } else {
                const uint32_t t = (X[0] ^ X[i]) & P;
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
    X[1] ^= X[0];
    X[2] ^= X[1];
    uint32_t t = 0;
    for (uint32_t Q = M; Q > 1; Q >>= 1)
        if (X[2] & Q) t ^= Q - 1;
    X[0] ^= t; X[1] ^= t; X[2] ^= t;
    uint32_t h = 0;
    for (int b = bits - 1; b >= 0; --b)
        for (int i = 0; i < 3; ++i) h = (h << 1) | ((X[i] >> b) & 1u);
    return h;
}

// --------------------------------------------------------------------------
// rebuild: binning
// --------------------------------------------------------------------------

// Start of a rebuild: clear the per-cell counters and extents, the per-system
// extents, the list status words, and mark every slot as padding.
__global__ void k_list_reset(int ncell, int *__restrict__ cnt, int *__restrict__ fill, unsigned *__restrict__ cellext,
                             int nsys, unsigned *__restrict__ sysext, int nstat, int *__restrict__ lst, int nslot,
                             int *__restrict__ perm, int *__restrict__ jcnt, int nmc, int *__restrict__ mcnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nmc) mcnt[i] = 0;
    if (i <= ncell) cnt[i] = 0;
    if (i < ncell) { fill[i] = 0; cellext[i] = 0u; }
    if (i < nsys) sysext[i] = 0u;
    if (i < nstat) lst[i] = 0;
    if (i < nslot) { perm[i] = -1; jcnt[i] = 0; }
}

// Half lists: an atom's listed partners are those of its own cluster's
// entries (icnt) and those it has as j in earlier clusters' entries (jcnt);
// md_run's limit applies to their sum (cutoff.f90:30-46).
__global__ void k_nbcheck(int nslot, const int *__restrict__ slot_sys, const SysP *__restrict__ sp,
                          const int *__restrict__ icnt, const int *__restrict__ jcnt, int *__restrict__ lst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nslot && icnt[i] + jcnt[i] > sp[slot_sys[i]].nblim) atomicOr(lst, 32);   // JP_ERR_NBLIM
}

// Wrap into the box and bin (celpar.f90:115-127 for md_run's cell numbering;
// the same wrapped coordinate g is binned on the box-aligned grid).
__global__ void k_bin(int natom, const int *__restrict__ atom_sys, const SysP *__restrict__ sp,
                      const double *__restrict__ x, int *__restrict__ img, int *__restrict__ cellid,
                      int *__restrict__ subkey, int *__restrict__ refcell, int *__restrict__ cnt,
                      int *__restrict__ err, const int *__restrict__ tflag, int *__restrict__ mcnt) {
#pragma clang fp contract(off)
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = a < natom;
    int cell = 0, mci = 0;
    bool mok = false;
    if (act) {
        const SysP &P = sp[atom_sys[a]];
        int c[3], m[3];
        for (int k = 0; k < 3; ++k) {
            double p = x[3 * a + k];
            double f = floor((p - P.wall[k]) * P.invL[k]);
            double g = p - P.L[k] * f;
            if (!(fabs(p) < 1e15)) { atomicOr(err, 1); f = 0; g = P.wall[k]; }
            img[3 * a + k] = (int)f;
            m[k] = (int)((g - P.rorg[k]) * P.rr) + 1;
            refcell[3 * a + k] = m[k];
            int ck = (int)floor((g - P.wall[k]) * P.icsz[k] * 2.0);   // half-cell resolution
            ck = ck < 0 ? 0 : (ck >= 2 * P.nc[k] ? 2 * P.nc[k] - 1 : ck);
            c[k] = ck;
        }
        // md_run's checks (celpar.f90:176-204): an atom outside its cell grid,
        // or more than ipmax atoms of one cluster class in a cell, is ier 28
        if (m[0] < 1 || m[0] > P.mnv || m[1] < 1 || m[1] > P.mnv || m[2] < 1 || m[2] > P.mnv) {
            atomicOr(err, 1);
        } else {
            const int lin = ((m[2] - 1) * P.mnv + (m[1] - 1)) * P.mnv + (m[0] - 1);
            mci = P.mc0 + 2 * lin + ((tflag[a] >> 13) & 1);
            mok = true;
        }
        // cells in Hilbert order; inside a cell, atoms ordered by sub-cell
        const uint32_t h = hilbert3((uint32_t)c[0], (uint32_t)c[1], (uint32_t)c[2], P.hbits + 1);
        cell = P.cell0 + (int)(h >> 3);
        cellid[a] = cell;
        subkey[a] = (int)(h & 7u);
    }
    const int old = wave_inc(mcnt, mci, mok);
    if (mok && old + 1 > sp[atom_sys[a]].ipmax) atomicOr(err, 1);
    (void)wave_inc(cnt, cell, act);
}

// Exclusive scan of n ints by one 1024-thread block (wave scans + block
// carry).  out[n] receives the total.
__global__ void __launch_bounds__(1024) k_scan(int n, const int *__restrict__ in, int *__restrict__ out) {
    __shared__ int wsum[16];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int base = 0; base < n; base += 1024) {
        int i = base + threadIdx.x;
        int v = i < n ? in[i] : 0;
        int s = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            int t = __shfl_up(s, o, 64);
            if (lane >= o) s += t;
        }
        if (lane == 63) wsum[w] = s;
        __syncthreads();
        if (w == 0) {
            int t = lane < 16 ? wsum[lane] : 0;
            int u = t;
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                int q = __shfl_up(u, o, 64);
                if (lane >= o) u += q;
            }
            if (lane < 16) wsum[lane] = u - t;
        }
        __syncthreads();
        int excl = carry + wsum[w] + s - v;
        if (i < n) out[i] = excl;
        __syncthreads();
        if (threadIdx.x == 1023) carry = excl + v;
        __syncthreads();
    }
    if (threadIdx.x == 0) out[n] = carry;
}

// Atoms of each cell at [start, start + cnt) as 64-bit keys (sub-cell << 32 |
// atom), in arrival order.
__global__ void k_place(int natom, const int *__restrict__ cellid, const int *__restrict__ subkey,
                        const int *__restrict__ start, int *__restrict__ fill, long long *__restrict__ tmp) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = a < natom;
    const int c = act ? cellid[a] : 0;
    const int p = wave_inc(fill, c, act);
    if (act) tmp[start[c] + p] = ((long long)subkey[a] << 32) | a;
}

// Deterministic order inside a cell: rank by (sub-cell, atom index).
__global__ void k_rank(int natom, const int *__restrict__ atom_sys, const SysP *__restrict__ sp,
                       const int *__restrict__ cellid, const int *__restrict__ subkey, const int *__restrict__ start,
                       const int *__restrict__ cnt, const long long *__restrict__ tmp, int *__restrict__ inum,
                       int *__restrict__ perm) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natom) return;
    const int c = cellid[a], b = start[c], n = cnt[c];
    const long long ka = ((long long)subkey[a] << 32) | a;
    int r = 0;
#pragma unroll 8
    for (int k = 0; k < n; ++k) r += tmp[b + k] < ka;
    const SysP &P = sp[atom_sys[a]];
    int slot = b + r - P.atom0 + P.slot0;
    inum[a] = slot;
    perm[slot] = a;
}

__global__ void k_cslot(int nc, const int *__restrict__ cs, const SysP *__restrict__ sp,
                        const int *__restrict__ start, int *__restrict__ cslot) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const SysP &P = sp[cs[c]];
    cslot[c] = start[c] - P.atom0 + P.slot0;
}

// Sorted copies for the list build (raw x, image) and the pair kernels (y).
__global__ void k_fill_sorted(int nslot, const int *__restrict__ slot_sys, const SysP *__restrict__ sp,
                              const int *__restrict__ perm, const double *__restrict__ x,
                              const int *__restrict__ img, const double *__restrict__ chg,
                              const int *__restrict__ tflag, const int *__restrict__ mol, JAtom *__restrict__ ys,
                              double *__restrict__ qd, double4 *__restrict__ xr, int4 *__restrict__ imgs,
                              float4 *__restrict__ lp, const int *__restrict__ refcell, int *__restrict__ mcell) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslot) return;
    int a = perm[s];
    const SysP &P = sp[slot_sys[s]];
    JAtom j;
    // md_run cell (celpar.f90:117-122) packed 8 bits per axis, for its neighbour-cell table
    mcell[s] = a < 0 ? 0
                     : (refcell[3 * a] & 0xff) | ((refcell[3 * a + 1] & 0xff) << 8) | ((refcell[3 * a + 2] & 0xff) << 16);
    if (a < 0) {
        j.x = P.wall[0] - 1.0e4; j.y = P.wall[1] - 1.0e4; j.z = P.wall[2] - 1.0e4;
        j.q = 0.f; j.t = 0;
        ys[s] = j;
        qd[s] = 0.0;
        xr[s] = make_double4(j.x, j.y, j.z, 0.0);
        imgs[s] = make_int4(0, 0, 0, -1);
        lp[s] = make_float4(-1.0e6f, -1.0e6f, -1.0e6f, __int_as_float(-1));
        return;
    }
    double p0 = x[3 * a], p1 = x[3 * a + 1], p2 = x[3 * a + 2];
    int i0 = img[3 * a], i1 = img[3 * a + 1], i2 = img[3 * a + 2];
    j.x = p0 - P.L[0] * (double)i0;
    j.y = p1 - P.L[1] * (double)i1;
    j.z = p2 - P.L[2] * (double)i2;
    j.q = (float)chg[a];
    j.t = tflag[a];
    ys[s] = j;
    qd[s] = chg[a];
    xr[s] = make_double4(p0, p1, p2, 0.0);
    imgs[s] = make_int4(i0, i1, i2, a);
    const int meta = (mol[a] & 0x3fffffff) | (((tflag[a] >> 12) & 1) << 30);
    lp[s] = make_float4((float)(j.x - P.ctr[0]), (float)(j.y - P.ctr[1]), (float)(j.z - P.ctr[2]), __int_as_float(meta));
}

// --------------------------------------------------------------------------
// per step: integration (verlet.f90:144-215) with the ZD excess term
// --------------------------------------------------------------------------

// Excess (excluded-pair) gradient of atom a, component k (enear.f90:528-550),
// FP64, in md_run's accumulation order, scaled per energy class.  Called from
// the pair kernels' epilogue, where no kernel is writing x.
__device__ __forceinline__ double excess_grad_k(int a, int k, const SysP &P, const double *__restrict__ x,
                                                const double *__restrict__ chg, const int *__restrict__ eptr,
                                                const int2 *__restrict__ eidx, const double *__restrict__ efac) {
#pragma clang fp contract(off)
    double g = 0.0;
    for (int e = eptr[a]; e < eptr[a + 1]; ++e) {
        const int2 pe = eidx[e];   // partner, class
        const int b = pe.x;
        double d = x[3 * a + k] - x[3 * b + k];
        d = d - P.L[k] * anint_d(d * P.invL[k]);
        double cc = chg[a] * chg[b];
        cc = cc * P.fcoeff;
        g = g + d * cc * efac[pe.y];
    }
    return g;
}

// Leap-frog kick and drift from the full gradient (excess + pair, per slot,
// written by the pair kernel).  With `gout` the gradient is also stored in
// original order for export; do_step == 0 only exports.
__global__ void k_kick(int natom, const int *__restrict__ atom_sys, const SysP *__restrict__ sp,
                       const int *__restrict__ inum, const double *__restrict__ gslot,
                       const double *__restrict__ mass, double *__restrict__ x, double *__restrict__ v,
                       double *__restrict__ xold, double *__restrict__ vback, double *__restrict__ gout, int do_step) {
#pragma clang fp contract(off)
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natom) return;
    const SysP &P = sp[atom_sys[a]];
    const size_t sl = inum[a];
    const double g0 = gslot[3 * sl], g1 = gslot[3 * sl + 1], g2 = gslot[3 * sl + 2];
    if (gout) { gout[3 * a] = g0; gout[3 * a + 1] = g1; gout[3 * a + 2] = g2; }
    if (!do_step) return;
    const double rt = -P.dt * 4.184e-4;   // md.f90:56 fdcvcv
    const double m = mass[a];
    double v0 = v[3 * a], v1 = v[3 * a + 1], v2 = v[3 * a + 2];
    if (vback) { vback[3 * a] = v0; vback[3 * a + 1] = v1; vback[3 * a + 2] = v2; }
    v0 = v0 + rt * g0 / m;
    v1 = v1 + rt * g1 / m;
    v2 = v2 + rt * g2 / m;
    double p0 = x[3 * a], p1 = x[3 * a + 1], p2 = x[3 * a + 2];
    xold[3 * a] = p0; xold[3 * a + 1] = p1; xold[3 * a + 2] = p2;
    x[3 * a] = p0 + P.dt * v0;
    x[3 * a + 1] = p1 + P.dt * v1;
    x[3 * a + 2] = p2 + P.dt * v2;
    v[3 * a] = v0; v[3 * a + 1] = v1; v[3 * a + 2] = v2;
}

// Excess energy per atom (half of each pair), classes separated; reduced on
// the host in atom order.
__global__ void k_excess_energy(int natom, const int *__restrict__ atom_sys, const SysP *__restrict__ sp,
                                const double *__restrict__ x, const double *__restrict__ chg,
                                const int *__restrict__ eptr, const int2 *__restrict__ eidx,
                                double *__restrict__ eout) {
#pragma clang fp contract(off)
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natom) return;
    const SysP &P = sp[atom_sys[a]];
    double e[3] = {0.0, 0.0, 0.0};
    for (int q = eptr[a]; q < eptr[a + 1]; ++q) {
        const int2 pe = eidx[q];
        const int b = pe.x;
        double dx = x[3 * a] - x[3 * b], dy = x[3 * a + 1] - x[3 * b + 1], dz = x[3 * a + 2] - x[3 * b + 2];
        dx = dx - P.L[0] * anint_d(dx * P.invL[0]);
        dy = dy - P.L[1] * anint_d(dy * P.invL[1]);
        dz = dz - P.L[2] * anint_d(dz * P.invL[2]);
        double dd2 = dx * dx + dy * dy + dz * dz;
        double cc = chg[a] * chg[b];
        double t = cc * P.bcoeff * dd2 * 0.5;
        e[0] += pe.y == 0 ? t : 0.0; e[1] += pe.y == 1 ? t : 0.0; e[2] += pe.y == 2 ? t : 0.0;
    }
    eout[3 * a] = e[0]; eout[3 * a + 1] = e[1]; eout[3 * a + 2] = e[2];
}

// --------------------------------------------------------------------------
// SHAKE (shake.f90), FP64, one thread per constraint group
// --------------------------------------------------------------------------

__device__ __forceinline__ void mimg3(const SysP &P, double *d) {
#pragma clang fp contract(off)
    for (int k = 0; k < 3; ++k) d[k] = d[k] - P.L[k] * anint_d(d[k] * P.invL[k]);
}

// shak2e_PB, shake.f90:687-755 (independent 2-atom groups, exact solution)
// shak2e_PB (shake.f90:687-755) on registers: r1 = the pair at t+dt (in/out),
// r2 = at t.  Exact solution of the single constraint.
__device__ __forceinline__ void shake2_regs(const SysP &P, double im1, double im2, double d2target,
                                            double (*r1)[3], const double (*r2)[3]) {
#pragma clang fp contract(off)
    const double im3 = im1 + im2, im4 = im3 * im3;
    double d1[3], d2[3];
    for (int k = 0; k < 3; ++k) {
        d1[k] = r1[1][k] - r1[0][k];
        d2[k] = r2[1][k] - r2[0][k];
    }
    mimg3(P, d1);
    mimg3(P, d2);
    double d1d1 = d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2];
    double d1d2 = d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2];
    double d2d2 = d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2];
    double A = im4 * d2d2, B = im3 * d1d2, C = d1d1 - d2target;
    double G = (B - sqrt(B * B - C * A)) / A;
    for (int k = 0; k < 3; ++k) {
        r1[0][k] = r1[0][k] + (im1 * G * d2[k]);
        r1[1][k] = r1[1][k] - (im2 * G * d2[k]);
    }
}

// shake3_PB (shake.f90:113-327) on registers: constraints (1,2), (2,3), (3,1),
// Newton iterations on the 3x3 system through its adjugate.  Returns 0, 4
// (not converged after maxit) or 8 (singular).
__device__ __forceinline__ int shake3_regs(const SysP &P, const double *im, const double *T, double tol, int maxit,
                                           double (*r1)[3], const double (*r2)[3]) {
#pragma clang fp contract(off)
    double co[3][3];
    co[0][0] = im[0] + im[1]; co[0][1] = -im[1]; co[0][2] = -im[0];
    co[1][0] = -im[1]; co[1][1] = im[1] + im[2]; co[1][2] = -im[2];
    co[2][0] = -im[0]; co[2][1] = -im[2]; co[2][2] = im[2] + im[0];
    const double TT[3] = {tol * T[0], tol * T[1], tol * T[2]};
    double G[6] = {0, 0, 0, 0, 0, 0};
    double d1[3][3], d2[3][3];
    for (int k = 0; k < 3; ++k) {
        d1[0][k] = r1[0][k] - r1[1][k];
        d1[1][k] = r1[1][k] - r1[2][k];
        d1[2][k] = r1[2][k] - r1[0][k];
        d2[0][k] = r2[0][k] - r2[1][k];
        d2[1][k] = r2[1][k] - r2[2][k];
        d2[2][k] = r2[2][k] - r2[0][k];
    }
    for (int q = 0; q < 3; ++q) { mimg3(P, d1[q]); mimg3(P, d2[q]); }
    for (int it = 0; it < maxit; ++it) {
        double B[3];
        for (int q = 0; q < 3; ++q) B[q] = T[q] - (d1[q][0] * d1[q][0] + d1[q][1] * d1[q][1] + d1[q][2] * d1[q][2]);
        if (fabs(B[0]) < TT[0] && fabs(B[1]) < TT[1] && fabs(B[2]) < TT[2]) {
            for (int k = 0; k < 3; ++k) {
                r1[0][k] = r1[0][k] + im[0] * (G[3] * d2[0][k] - G[5] * d2[2][k]);
                r1[1][k] = r1[1][k] + im[1] * (-G[3] * d2[0][k] + G[4] * d2[1][k]);
                r1[2][k] = r1[2][k] + im[2] * (-G[4] * d2[1][k] + G[5] * d2[2][k]);
            }
            return 0;
        }
        double gf[3][3];
        for (int ig = 0; ig < 3; ++ig)
            for (int q = 0; q < 3; ++q)
                gf[q][ig] = 2.0 * co[q][ig] * (d1[q][0] * d2[ig][0] + d1[q][1] * d2[ig][1] + d1[q][2] * d2[ig][2]);
        double deta = (gf[0][0] * gf[1][1] * gf[2][2]) + (gf[0][1] * gf[1][2] * gf[2][0]) +
                      (gf[0][2] * gf[1][0] * gf[2][1]) - (gf[0][0] * gf[1][2] * gf[2][1]) -
                      (gf[0][1] * gf[1][0] * gf[2][2]) - (gf[0][2] * gf[1][1] * gf[2][0]);
        if (deta == 0.0) return 8;
        deta = 1.0 / deta;
        double db[3][3];
        db[0][0] = gf[1][1] * gf[2][2] - gf[1][2] * gf[2][1];
        db[0][1] = gf[2][1] * gf[0][2] - gf[2][2] * gf[0][1];
        db[0][2] = gf[0][1] * gf[1][2] - gf[0][2] * gf[1][1];
        db[1][0] = gf[1][2] * gf[2][0] - gf[1][0] * gf[2][2];
        db[1][1] = gf[2][2] * gf[0][0] - gf[2][0] * gf[0][2];
        db[1][2] = gf[0][2] * gf[1][0] - gf[0][0] * gf[1][2];
        db[2][0] = gf[1][0] * gf[2][1] - gf[1][1] * gf[2][0];
        db[2][1] = gf[2][0] * gf[0][1] - gf[2][1] * gf[0][0];
        db[2][2] = gf[0][0] * gf[1][1] - gf[0][1] * gf[1][0];
        double Gn[3];
        for (int q = 0; q < 3; ++q) Gn[q] = deta * (db[q][0] * B[0] + db[q][1] * B[1] + db[q][2] * B[2]);
        for (int q = 0; q < 3; ++q) { G[q] = Gn[q]; G[3 + q] = G[3 + q] + Gn[q]; }
        for (int q = 0; q < 3; ++q)
            for (int k = 0; k < 3; ++k)
                d1[q][k] = d1[q][k] + co[q][0] * d2[0][k] * G[0] + co[q][1] * d2[1][k] * G[1] +
                           co[q][2] * d2[2][k] * G[2];
    }
    return 4;
}

// shake4_PB (shake.f90:333-562) on one 4-atom group held in registers: six
// constraints, pairs (1,2) (2,3) (3,1) (4,1) (3,4) (2,4) (parlis), Newton
// iterations whose 6 x 6 systems are solved as simulu does (simulu.f90:
// ludeco's partial pivoting by row indices, pivots below eps^3 = 1e-30
// singular, then the two triangular solves; row-major a[i][j] = A(i, j)).
// Returns 0, 4 (not converged: r1 unchanged) or 8 (singular).
__device__ __forceinline__ int shake4_regs(const SysP &P, const double *im, const double *T, double tol, int maxit,
                                           double (*r1)[3], const double (*r2)[3]) {
#pragma clang fp contract(off)
    const double co[6][6] = {{im[0] + im[1], -im[1], -im[0], -im[0], 0.0, -im[1]},
                             {-im[1], im[1] + im[2], -im[2], 0.0, -im[2], im[1]},
                             {-im[0], -im[2], im[2] + im[0], im[0], im[2], 0.0},
                             {-im[0], 0.0, im[0], im[3] + im[0], -im[3], -im[3]},
                             {0.0, -im[2], im[2], -im[3], im[2] + im[3], im[3]},
                             {-im[1], im[1], 0.0, -im[3], im[3], im[1] + im[3]}};
    const int p0[6] = {0, 1, 2, 3, 2, 1}, p1[6] = {1, 2, 0, 0, 3, 3};
    double d1[6][3], d2[6][3], w[6][3], G[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        G[q] = 0.0;
        for (int k = 0; k < 3; ++k) {
            d1[q][k] = r1[p0[q]][k] - r1[p1[q]][k];
            d2[q][k] = r2[p0[q]][k] - r2[p1[q]][k];
        }
        mimg3(P, d1[q]);
        mimg3(P, d2[q]);
        for (int k = 0; k < 3; ++k) w[q][k] = d1[q][k];
    }
    for (int it = 0; it < maxit; ++it) {
        double B[6];
        bool ok = true;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            B[q] = T[q] - (w[q][0] * w[q][0] + w[q][1] * w[q][1] + w[q][2] * w[q][2]);
            if (fabs(B[q]) >= tol * T[q]) ok = false;
        }
        if (ok) {
            for (int k = 0; k < 3; ++k) {
                r1[0][k] = r1[0][k] + im[0] * (G[0] * d2[0][k] - G[2] * d2[2][k] - G[3] * d2[3][k]);
                r1[1][k] = r1[1][k] + im[1] * (-G[0] * d2[0][k] + G[1] * d2[1][k] + G[5] * d2[5][k]);
                r1[2][k] = r1[2][k] + im[2] * (-G[1] * d2[1][k] + G[2] * d2[2][k] + G[4] * d2[4][k]);
                r1[3][k] = r1[3][k] + im[3] * (G[3] * d2[3][k] - G[4] * d2[4][k] - G[5] * d2[5][k]);
            }
            return 0;
        }
        double a[6][6];
#pragma unroll
        for (int ig = 0; ig < 6; ++ig)
#pragma unroll
            for (int q = 0; q < 6; ++q)
                a[q][ig] = 2.0 * co[q][ig] * (w[q][0] * d2[ig][0] + w[q][1] * d2[ig][1] + w[q][2] * d2[ig][2]);
#pragma unroll
        for (int ig = 0; ig < 6; ++ig)
#pragma unroll
            for (int q = 0; q < 6; ++q) B[q] = B[q] + a[q][ig] * G[ig];
        // ludeco: the permutation as row indices; rows are addressed through
        // ip, kept in registers by selecting with unrolled compares
        int ip[6] = {0, 1, 2, 3, 4, 5};
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            int l = k;
            double al = 0.0;
#pragma unroll
            for (int r = 0; r < 6; ++r)
                if (r == ip[k]) al = fabs(a[r][k]);
#pragma unroll
            for (int i = k + 1; i < 6; ++i) {
                double v = 0.0;
#pragma unroll
                for (int r = 0; r < 6; ++r)
                    if (r == ip[i]) v = fabs(a[r][k]);
                if (v > al) { l = i; al = v; }
            }
            if (l != k) { const int t = ip[k]; ip[k] = ip[l]; ip[l] = t; }
            if (al < 1.0e-30) return 8;   // PHYCNS.f90:34 eps = 1e-10, cubed
            double piv[6];
#pragma unroll
            for (int j = 0; j < 6; ++j) {
                piv[j] = 0.0;
#pragma unroll
                for (int r = 0; r < 6; ++r)
                    if (r == ip[k]) piv[j] = a[r][j];
            }
            piv[k] = 1.0 / piv[k];
#pragma unroll
            for (int r = 0; r < 6; ++r)
                if (r == ip[k]) a[r][k] = piv[k];
#pragma unroll
            for (int i = k + 1; i < 6; ++i) {
                const int ri = ip[i];
#pragma unroll
                for (int r = 0; r < 6; ++r) {
                    if (r != ri) continue;
                    a[r][k] = a[r][k] * piv[k];
#pragma unroll
                    for (int j = k + 1; j < 6; ++j) a[r][j] = a[r][j] - a[r][k] * piv[j];
                }
            }
        }
        double x[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            double t = 0.0, row[6];
#pragma unroll
            for (int r = 0; r < 6; ++r)
                if (r == ip[i]) {
                    t = B[r];
#pragma unroll
                    for (int j = 0; j < 6; ++j) row[j] = a[r][j];
                }
#pragma unroll
            for (int j = 0; j < i; ++j) t = t - row[j] * x[j];
            x[i] = t;
        }
#pragma unroll
        for (int i = 5; i >= 0; --i) {
            double row[6];
#pragma unroll
            for (int r = 0; r < 6; ++r)
                if (r == ip[i])
#pragma unroll
                    for (int j = 0; j < 6; ++j) row[j] = a[r][j];
            double t = x[i];
#pragma unroll
            for (int j = i + 1; j < 6; ++j) t = t - row[j] * x[j];
            x[i] = t * row[i];
        }
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            G[q] = x[q];
            for (int k = 0; k < 3; ++k) w[q][k] = d1[q][k];
        }
#pragma unroll
        for (int ig = 0; ig < 6; ++ig)
#pragma unroll
            for (int q = 0; q < 6; ++q)
                for (int k = 0; k < 3; ++k) w[q][k] = w[q][k] + co[q][ig] * d2[ig][k] * G[ig];
    }
    return 4;
}

__global__ void __launch_bounds__(128) k_shake4(int ng, const int *__restrict__ gsys, const SysP *__restrict__ sp,
                                                const int4 *__restrict__ ga, const double *__restrict__ gd,
                                                const double *__restrict__ imass, double *__restrict__ r1,
                                                const double *__restrict__ r2, double tol, int maxit,
                                                int *__restrict__ err) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    const SysP &P = sp[gsys[g]];
    const int a[4] = {ga[g].x, ga[g].y, ga[g].z, ga[g].w};
    double im[4], T[6], x1[4][3], x2[4][3];
    for (int q = 0; q < 4; ++q) {
        im[q] = imass[a[q]];
        for (int k = 0; k < 3; ++k) { x1[q][k] = r1[3 * a[q] + k]; x2[q][k] = r2[3 * a[q] + k]; }
    }
    for (int q = 0; q < 6; ++q) T[q] = gd[6 * g + q];
    const int e = shake4_regs(P, im, T, tol, maxit, x1, x2);
    if (e) { atomicOr(err, e); return; }
    for (int q = 0; q < 4; ++q)
        for (int k = 0; k < 3; ++k) r1[3 * a[q] + k] = x1[q][k];
}

__global__ void __launch_bounds__(128) k_shake2(int ng, const int *__restrict__ gsys, const SysP *__restrict__ sp,
                         const int4 *__restrict__ ga, const double *__restrict__ gd, const double *__restrict__ mass,
                         double *__restrict__ r1, const double *__restrict__ r2) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    const SysP &P = sp[gsys[g]];
    const int a[2] = {ga[g].x, ga[g].y};
    double x1[2][3], x2[2][3];
    for (int q = 0; q < 2; ++q)
        for (int k = 0; k < 3; ++k) { x1[q][k] = r1[3 * a[q] + k]; x2[q][k] = r2[3 * a[q] + k]; }
    shake2_regs(P, 1.0 / mass[a[0]], 1.0 / mass[a[1]], gd[6 * g], x1, x2);
    for (int q = 0; q < 2; ++q)
        for (int k = 0; k < 3; ++k) r1[3 * a[q] + k] = x1[q][k];
}

__global__ void __launch_bounds__(128) k_shake3(int ng, const int *__restrict__ gsys, const SysP *__restrict__ sp,
                         const int4 *__restrict__ ga, const double *__restrict__ gd,
                         const double *__restrict__ imass, double *__restrict__ r1, const double *__restrict__ r2,
                         double tol, int maxit, int *__restrict__ err) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    const SysP &P = sp[gsys[g]];
    const int a[3] = {ga[g].x, ga[g].y, ga[g].z};
    const double im[3] = {imass[a[0]], imass[a[1]], imass[a[2]]};
    const double T[3] = {gd[6 * g], gd[6 * g + 1], gd[6 * g + 2]};
    double x1[3][3], x2[3][3];
    for (int q = 0; q < 3; ++q)
        for (int k = 0; k < 3; ++k) { x1[q][k] = r1[3 * a[q] + k]; x2[q][k] = r2[3 * a[q] + k]; }
    const int e = shake3_regs(P, im, T, tol, maxit, x1, x2);
    if (e) { atomicOr(err, e); return; }
    for (int q = 0; q < 3; ++q)
        for (int k = 0; k < 3; ++k) r1[3 * a[q] + k] = x1[q][k];
}

// One MD step after the forces, fused per "unit" (a 3-atom SHAKE group, a
// 2-atom group or a free atom): leap-frog kick and drift (verlet.f90:144-190),
// SHAKE on the unit's atoms in registers, v = (x - x_old)/dt when the run
// has constraints (verlet.f90:196-215), and the refreshed sorted coordinates.
// The same operations in the same order as k_kick -> k_shake* -> k_post.
// unit = {a0, a1, a2, kind | param << 3}, kind 1-4 = atoms in the unit (the
// fourth atom of a 4-atom group in s4a[param].w).  NQ: 3, or 4 when the
// run has 4-atom groups.
// KICK false: v already holds the kicked velocity (a step that stops the
// centre-of-mass motion between kick and drift); VUPD false: v keeps it
// after SHAKE (md_run's initial step, staver.f90:196-204).
template <bool KICK, bool VUPD, int NQ>
__global__ void __launch_bounds__(128) k_step(int nunit, const int4 *__restrict__ unit, const int *__restrict__ atom_sys,
                       const SysP *__restrict__ sp, const int *__restrict__ inum, const int *__restrict__ img,
                       const double *__restrict__ gslot, const double *__restrict__ mass,
                       const double *__restrict__ imass, const double *__restrict__ s2d,
                       const double *__restrict__ s3d, const int4 *__restrict__ s4a, const double *__restrict__ s4d,
                       double tol, int maxit, int shake_on, double *__restrict__ x,
                       double *__restrict__ v, double *__restrict__ vback, JAtom *__restrict__ ys,
                       int *__restrict__ err) {
#pragma clang fp contract(off)
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nunit) return;
    const int4 U = unit[u];
    const int kind = U.w & 7, prm = U.w >> 3;
    int a[NQ];
    a[0] = U.x; a[1] = U.y; a[2] = U.z;
    if constexpr (NQ == 4) a[3] = kind == 4 ? s4a[prm].w : U.x;
    const SysP &P = sp[atom_sys[a[0]]];
    const double rt = -P.dt * 4.184e-4;   // md.f90:56 fdcvcv
    double xn[NQ][3], xo[NQ][3], vv[NQ][3];
    // what SHAKE and the sorted copy need, loaded up front so that these
    // loads overlap the first ones instead of following the SHAKE loop
    int sl3[NQ], im3[NQ][3];
    double m3[NQ], iw[NQ], T3[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int b = q < kind ? a[q] : a[0];
        sl3[q] = inum[b];
        m3[q] = mass[b];
        iw[q] = 0.0;
        for (int k = 0; k < 3; ++k) im3[q][k] = img[3 * b + k];
    }
    if (NQ == 4 && kind == 4) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) iw[q] = imass[a[q]];
#pragma unroll
        for (int q = 0; q < 6; ++q) T3[q] = s4d[6 * prm + q];
    } else if (kind == 3) {
        iw[0] = imass[a[0]]; iw[1] = imass[a[1]]; iw[2] = imass[a[2]];
        T3[0] = s3d[6 * prm]; T3[1] = s3d[6 * prm + 1]; T3[2] = s3d[6 * prm + 2];
    } else if (kind == 2) {
        T3[0] = s2d[6 * prm];
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {   // fixed trip count: the arrays stay in registers
        if (q >= kind) break;
        const int b = a[q];
        const size_t sl = sl3[q];
        const double m = m3[q];
        for (int k = 0; k < 3; ++k) {
            double vk = v[3 * b + k];
            if (KICK) {
                const double g = gslot[3 * sl + k];
                if (vback) vback[3 * b + k] = vk;
                vk = vk + rt * g / m;
            }
            const double p = x[3 * b + k];
            xo[q][k] = p;
            xn[q][k] = p + P.dt * vk;
            vv[q][k] = vk;
        }
    }
    if constexpr (NQ == 4) {
        if (kind == 4) {
            const int e = shake4_regs(P, iw, T3, tol, maxit, xn, xo);
            if (e) atomicOr(err, e);
        }
    }
    if (kind == 3) {
        const int e = shake3_regs(P, iw, T3, tol, maxit, xn, xo);
        if (e) atomicOr(err, e);
    } else if (kind == 2) {
        shake2_regs(P, 1.0 / m3[0], 1.0 / m3[1], T3[0], xn, xo);
    }
    const double r = 1.0 / P.dt;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        if (q >= kind) break;
        const int b = a[q];
        JAtom *j = ys + sl3[q];
        for (int k = 0; k < 3; ++k) {
            x[3 * b + k] = xn[q][k];
            if (VUPD) v[3 * b + k] = shake_on ? (xn[q][k] - xo[q][k]) * r : vv[q][k];
        }
        j->x = xn[q][0] - P.L[0] * (double)im3[q][0];
        j->y = xn[q][1] - P.L[1] * (double)im3[q][1];
        j->z = xn[q][2] - P.L[2] * (double)im3[q][2];
    }
}

__global__ void k_post(int natom, const int *__restrict__ atom_sys, const SysP *__restrict__ sp,
                       const int *__restrict__ inum, const int *__restrict__ img, const double *__restrict__ x,
                       const double *__restrict__ xold, double *__restrict__ v, JAtom *__restrict__ ys,
                       int shake_on) {
#pragma clang fp contract(off)
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natom) return;
    const SysP &P = sp[atom_sys[a]];
    double p0 = x[3 * a], p1 = x[3 * a + 1], p2 = x[3 * a + 2];
    if (shake_on) {
        double r = 1.0 / P.dt;
        v[3 * a] = (p0 - xold[3 * a]) * r;
        v[3 * a + 1] = (p1 - xold[3 * a + 1]) * r;
        v[3 * a + 2] = (p2 - xold[3 * a + 2]) * r;
    }
    JAtom *j = ys + inum[a];
    j->x = p0 - P.L[0] * (double)img[3 * a];
    j->y = p1 - P.L[1] * (double)img[3 * a + 1];
    j->z = p2 - P.L[2] * (double)img[3 * a + 2];
}

// Kinetic energy terms m|v|^2 with v = (vback + v)/2 (md.f90:618-628); summed
// per system on the host in atom order.
// Per-segment sums in a fixed order (deterministic run to run), two passes.
// k_seg_part: block (j, s) adds a[i * ncomp + k] over i = beg[s] + j*256 + t,
// + JP_SEG_NB*256, ... (thread t, in sequence), then a fixed tree over its 256
// threads -> part[(s * JP_SEG_NB + j) * ncomp + k].  k_seg_fin: the JP_SEG_NB
// partials of segment s in order -> out[s * ncomp + k].  ncomp <= 8.
#define JP_SEG_NB 64
template <typename T>
__global__ void __launch_bounds__(256) k_seg_part(const int *__restrict__ beg, const int *__restrict__ end, int ncomp,
                                                  const T *__restrict__ a, T *__restrict__ part) {
#pragma clang fp contract(off)
    __shared__ T s[8][256];
    const int j = blockIdx.x, sg = blockIdx.y, t = threadIdx.x;
    T acc[8];
    for (int k = 0; k < 8; ++k) acc[k] = T(0);
    for (int i = beg[sg] + j * 256 + t; i < end[sg]; i += JP_SEG_NB * 256)
        for (int k = 0; k < ncomp; ++k) acc[k] = acc[k] + a[(size_t)i * ncomp + k];
    for (int k = 0; k < ncomp; ++k) s[k][t] = acc[k];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h)
            for (int k = 0; k < ncomp; ++k) s[k][t] = s[k][t] + s[k][t + h];
        __syncthreads();
    }
    if (t < ncomp) part[((size_t)sg * JP_SEG_NB + j) * ncomp + t] = s[t][0];
}

template <typename T>
__global__ void k_seg_fin(int ncomp, const T *__restrict__ part, T *__restrict__ out) {
#pragma clang fp contract(off)
    const int sg = blockIdx.x, t = threadIdx.x;
    if (t >= ncomp) return;
    T acc = T(0);
    for (int j = 0; j < JP_SEG_NB; ++j) acc = acc + part[((size_t)sg * JP_SEG_NB + j) * ncomp + t];
    out[(size_t)sg * ncomp + t] = acc;
}

__global__ void k_ekin(int natom, const double *__restrict__ mass, const double *__restrict__ v,
                       const double *__restrict__ vback, double *__restrict__ out) {
#pragma clang fp contract(off)
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natom) return;
    double w0 = 0.5 * (vback[3 * a] + v[3 * a]);
    double w1 = 0.5 * (vback[3 * a + 1] + v[3 * a + 1]);
    double w2 = 0.5 * (vback[3 * a + 2] + v[3 * a + 2]);
    out[a] = mass[a] * (w0 * w0 + w1 * w1 + w2 * w2);
}
