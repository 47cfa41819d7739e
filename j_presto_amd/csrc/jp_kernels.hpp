// Device code of libjpgpu: MI355X (gfx950) kernels for j_presto md_run's
// ZD + LJ nonbonded path, the pair-list build that feeds it, and the FP64
// SHAKE / leap-frog step.
//
// Data layout in HBM (one or many systems concatenated; a single system is a
// batch of one):
//   original order  x, v, xold: double[N][3]; per-atom topology arrays
//   sorted "slots"  atoms ordered by (system, box-aligned cell in Morton
//                   order, atom index); every system's slot range is padded
//                   to a multiple of 64 so that 64 consecutive slots form one
//                   i-group = one wavefront.
//     ys[slot]      JAtom {y (FP64, coordinates wrapped at the last rebuild),
//                   q (FP32 chgmod), packed type/class/vdW flags} -- 32 B, one
//                   scalar load per j atom in the pair loop
//     qd[slot]      chgmod in FP64 (plain FP64 path)
//     xr[slot]      raw coordinates at the last rebuild (list test)
//   per group       a j-list of Entry {j slot | image shift, 64-bit lane mask}:
//                   bit l of the mask = (slot 64g+l, j) is in md_run's pair list
//                   (FP64 minimum-image d^2 <= rlim^2, not excluded).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define JP_WAVE 64
#define JP_FSCALE 1099511627776.0   /* 2^40: force fixed point, +-8.4e6 kcal/mol/A */
#define JP_ESCALE 68719476736.0     /* 2^36: energy fixed point, +-1.3e8 kcal/mol */
#define JP_SLOT_BITS 27
#define JP_SLOT_MASK ((1u << JP_SLOT_BITS) - 1u)
#define JP_FLUSH 32                 /* FP32 partial sums flushed to int64 every 32 j */

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
    int grp0, ngrp;                  // i-group range
    int lj0, ntyp;                   // LJ table offset / size
    int hbits;                       // nc = 2^hbits cells per edge
    int small;                       // some edge < 2*rlim: minimum image per pair
    double dt;
};

struct JAtom {          // 32 B
    double x, y, z;
    float q;
    int t;              // bits 0-11 type, bit 12 has-vdW, bit 13 cluster class
};

struct Entry {          // 16 B
    uint32_t js;        // j slot | shift code << 27 (code = (sx+1)*9+(sy+1)*3+(sz+1))
    uint32_t pad;
    uint64_t m;         // lane mask
};

__device__ __forceinline__ double anint_d(double x) {
    // Fortran ANINT (round half away from zero); x - trunc(x) is exact
    double t = trunc(x);
    double r = x - t;
    return (fabs(r) >= 0.5) ? t + copysign(1.0, x) : t;
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & (JP_WAVE - 1); }

// XCD-aware block -> work-item mapping: the 8 XCDs take blocks round-robin;
// give each XCD a contiguous range so neighbouring groups share its L2.
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

// Wrap into the box and bin (celpar.f90:115-127 for md_run's cell numbering;
// the same wrapped coordinate g is binned on the box-aligned grid).
__global__ void k_bin(int natom, const int *__restrict__ atom_sys, const SysP *__restrict__ sp,
                      const double *__restrict__ x, int *__restrict__ img, int *__restrict__ cellid,
                      int *__restrict__ subkey, int *__restrict__ refcell, int *__restrict__ cnt,
                      int *__restrict__ err) {
#pragma clang fp contract(off)
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natom) return;
    const SysP &P = sp[atom_sys[a]];
    int c[3];
    for (int k = 0; k < 3; ++k) {
        double p = x[3 * a + k];
        double f = floor((p - P.wall[k]) * P.invL[k]);
        double g = p - P.L[k] * f;
        if (!(fabs(p) < 1e15)) { atomicOr(err, 1); f = 0; g = P.wall[k]; }
        img[3 * a + k] = (int)f;
        refcell[3 * a + k] = (int)((g - P.rorg[k]) * P.rr) + 1;
        int ck = (int)floor((g - P.wall[k]) * P.icsz[k] * 2.0);   // half-cell resolution
        ck = ck < 0 ? 0 : (ck >= 2 * P.nc[k] ? 2 * P.nc[k] - 1 : ck);
        c[k] = ck;
    }
    // cells in Hilbert order; inside a cell, atoms ordered by sub-cell
    const uint32_t h = hilbert3((uint32_t)c[0], (uint32_t)c[1], (uint32_t)c[2], P.hbits + 1);
    int cell = P.cell0 + (int)(h >> 3);
    cellid[a] = cell;
    subkey[a] = (int)(h & 7u);
    atomicAdd(&cnt[cell], 1);
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

__global__ void k_place(int natom, const int *__restrict__ cellid, const int *__restrict__ start,
                        int *__restrict__ fill, int *__restrict__ tmp) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natom) return;
    int c = cellid[a];
    int p = atomicAdd(&fill[c], 1);
    tmp[start[c] + p] = a;
}

// Deterministic order inside a cell: rank by (sub-cell, atom index).
__global__ void k_rank(int natom, const int *__restrict__ atom_sys, const SysP *__restrict__ sp,
                       const int *__restrict__ cellid, const int *__restrict__ subkey, const int *__restrict__ start,
                       const int *__restrict__ cnt, const int *__restrict__ tmp, int *__restrict__ inum,
                       int *__restrict__ perm) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natom) return;
    int c = cellid[a], b = start[c], n = cnt[c], r = 0;
    const long long ka = ((long long)subkey[a] << 32) | a;
    for (int k = 0; k < n; ++k) {
        const int o = tmp[b + k];
        r += (((long long)subkey[o] << 32) | o) < ka;
    }
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
                              const int *__restrict__ tflag, JAtom *__restrict__ ys, double *__restrict__ qd,
                              double4 *__restrict__ xr, int4 *__restrict__ imgs) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslot) return;
    int a = perm[s];
    const SysP &P = sp[slot_sys[s]];
    JAtom j;
    if (a < 0) {
        j.x = P.wall[0] - 1.0e4; j.y = P.wall[1] - 1.0e4; j.z = P.wall[2] - 1.0e4;
        j.q = 0.f; j.t = 0;
        ys[s] = j;
        qd[s] = 0.0;
        xr[s] = make_double4(j.x, j.y, j.z, 0.0);
        imgs[s] = make_int4(0, 0, 0, -1);
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
}

// Bounding box of each group's wrapped coordinates.
__global__ void __launch_bounds__(64) k_bbox(int ngrp, const int *__restrict__ gcnt, const JAtom *__restrict__ ys,
                                             double *__restrict__ bb) {
    int g = blockIdx.x;
    if (g >= ngrp) return;
    int l = threadIdx.x;
    bool act = l < gcnt[g];
    JAtom j = ys[(size_t)g * 64 + l];
    double lo[3] = {act ? j.x : 1e300, act ? j.y : 1e300, act ? j.z : 1e300};
    double hi[3] = {act ? j.x : -1e300, act ? j.y : -1e300, act ? j.z : -1e300};
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        for (int k = 0; k < 3; ++k) {
            lo[k] = fmin(lo[k], __shfl_xor(lo[k], o, 64));
            hi[k] = fmax(hi[k], __shfl_xor(hi[k], o, 64));
        }
    if (l == 0)
        for (int k = 0; k < 3; ++k) { bb[6 * g + k] = lo[k]; bb[6 * g + 3 + k] = hi[k]; }
}

// --------------------------------------------------------------------------
// rebuild: the pair list (cutoff.f90 cutatm_PB + cutoff_PB semantics)
// --------------------------------------------------------------------------
//
// One block of JP_WL wavefronts per i-group; wave w takes every JP_WL-th
// candidate cell.  Candidate j atoms: every atom of every cell of the
// unwrapped cell range around the group's bounding box grown by rlim, each
// with the periodic shift of that unwrapped cell.  Lane l sets its bit for
// (j, shift) iff
//   (a) md_run's FP64 test passes on raw coordinates: d = xi - xj,
//       d -= L*anint(d/L), dx*dx+dy*dy+dz*dz <= rlim^2 (cutoff.f90:464-469),
//   (b) the shift is that pair's minimum image (each pair appears once),
//   (c) j != i and (i,j) is not an excluded pair (cutoff.f90:50-96).
// Pipeline: k_list_cap (entries can not exceed candidates) -> scan ->
// k_list (one pass; a cell chunk is loaded coalesced, every lane tests its
// own j against the group box first, only survivors get the 64-lane test)
// -> scan -> k_list_compact.
#define JP_WL 4

struct CandRange {
    int ulo[3], n[3], ncand;
    double lo[3], hi[3], rl;
    bool bad;
};

__device__ __forceinline__ CandRange cand_range(const SysP &P, const double *__restrict__ bb, int g) {
    CandRange c;
    c.rl = sqrt(P.rlim2) * (1.0 + 1e-12) + 1e-9;  // conservative pruning radius
    c.bad = false;
    for (int k = 0; k < 3; ++k) {
        c.lo[k] = bb[6 * g + k];
        c.hi[k] = bb[6 * g + 3 + k];
        double a = floor((c.lo[k] - c.rl - P.wall[k]) * P.icsz[k]);
        double b = floor((c.hi[k] + c.rl - P.wall[k]) * P.icsz[k]);
        // non-finite or runaway coordinates: report and do nothing (bounded loops)
        if (!(a >= -4.0 * P.nc[k] - 64.0) || !(b <= 5.0 * P.nc[k] + 64.0) || !(b - a <= 3.0 * P.nc[k] + 16.0))
            c.bad = true;
        c.ulo[k] = c.bad ? 0 : (int)a;
        c.n[k] = c.bad ? 0 : (int)(b - a) + 1;
    }
    c.ncand = c.bad ? 0 : c.n[0] * c.n[1] * c.n[2];
    return c;
}

// Candidate q -> (cell, shift code) or -1 when the unwrapped cell is farther
// than rlim from the group box.
__device__ __forceinline__ int cand_cell(const SysP &P, const CandRange &c, const int *__restrict__ crank, int q,
                                         int &sx, int &sy, int &sz) {
    const int ux = c.ulo[0] + q % c.n[0], uy = c.ulo[1] + (q / c.n[0]) % c.n[1], uz = c.ulo[2] + q / (c.n[0] * c.n[1]);
    double gx = fmax(0.0, fmax(P.wall[0] + ux * P.csz[0] - c.hi[0], c.lo[0] - (P.wall[0] + (ux + 1) * P.csz[0])));
    double gy = fmax(0.0, fmax(P.wall[1] + uy * P.csz[1] - c.hi[1], c.lo[1] - (P.wall[1] + (uy + 1) * P.csz[1])));
    double gz = fmax(0.0, fmax(P.wall[2] + uz * P.csz[2] - c.hi[2], c.lo[2] - (P.wall[2] + (uz + 1) * P.csz[2])));
    if (gx * gx + gy * gy + gz * gz > c.rl * c.rl) return -1;
    sx = (ux >= 0) ? ux / P.nc[0] : -((-ux + P.nc[0] - 1) / P.nc[0]);
    sy = (uy >= 0) ? uy / P.nc[1] : -((-uy + P.nc[1] - 1) / P.nc[1]);
    sz = (uz >= 0) ? uz / P.nc[2] : -((-uz + P.nc[2] - 1) / P.nc[2]);
    const int cx = ux - sx * P.nc[0], cy = uy - sy * P.nc[1], cz = uz - sz * P.nc[2];
    return crank[P.cell0 + (cz * P.nc[1] + cy) * P.nc[0] + cx];
}

// Upper bound of each (group, wave)'s entries: its candidate atoms.
__global__ void __launch_bounds__(64)
k_list_cap(int ngrp, const int *__restrict__ grp_sys, const SysP *__restrict__ sp, const double *__restrict__ bb,
           const int *__restrict__ crank, const int *__restrict__ ccnt, int *__restrict__ cap,
           int *__restrict__ err) {
    const int gw = blockIdx.x * 64 + threadIdx.x;   // one thread per (group, wave)
    if (gw >= ngrp * JP_WL) return;
    const int g = gw / JP_WL, w = gw % JP_WL;
    const SysP &P = sp[grp_sys[g]];
    const CandRange c = cand_range(P, bb, g);
    if (c.bad) atomicOr(err, 1);
    int n = 0;
    for (int q = w; q < c.ncand; q += JP_WL) {
        int sx, sy, sz;
        int cell = cand_cell(P, c, crank, q, sx, sy, sz);
        if (cell >= 0) n += ccnt[cell];
    }
    cap[gw] = n;
}

__device__ __forceinline__ double bcast_d(double v, int k) {
    int lo = __builtin_amdgcn_readlane(__double2loint(v), k);
    int hi = __builtin_amdgcn_readlane(__double2hiint(v), k);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ int bcast_i(int v, int k) { return __builtin_amdgcn_readlane(v, k); }

__global__ void __launch_bounds__(64 * JP_WL)
k_list(int ngrp, const int *__restrict__ grp_sys, const int *__restrict__ gcnt, const SysP *__restrict__ sp,
       const double *__restrict__ bb, const int *__restrict__ crank, const int *__restrict__ cslot,
       const int *__restrict__ ccnt, const double4 *__restrict__ xr, const int4 *__restrict__ imgs,
       const int *__restrict__ mol, const int *__restrict__ xptr, const int *__restrict__ xidx,
       const int *__restrict__ coff, Entry *__restrict__ tmp, int *__restrict__ lcount,
       int *__restrict__ nremoved, int *__restrict__ nbcount) {
#pragma clang fp contract(off)
    __shared__ int s_nb[JP_WL][64];
    __shared__ int s_nrem[JP_WL];
    int g = xcd_remap(blockIdx.x, gridDim.x);
    if (g >= ngrp) return;
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const SysP &P = sp[grp_sys[g]];
    const bool act = l < gcnt[g];
    const int slot = g * 64 + l;
    const double4 xi = xr[slot];
    const int4 ii = imgs[slot];
    const int ai = act ? ii.w : -1;
    const int moli = act ? mol[ai] : -2;
    const int xb = act ? xptr[ai] : 0, xe = act ? xptr[ai + 1] : 0;
    const CandRange c = cand_range(P, bb, g);
    const double rl2 = c.rl * c.rl;
    const int pos0 = coff[g * JP_WL + w];
    int pos = pos0;
    int nrem = 0, nb = 0;
    for (int q = w; q < c.ncand; q += JP_WL) {
        int sx, sy, sz;
        const int cell = cand_cell(P, c, crank, q, sx, sy, sz);
        if (cell < 0) continue;
        const int j0 = cslot[cell], nj = ccnt[cell];
        const uint32_t code = (uint32_t)((sx + 1) * 9 + (sy + 1) * 3 + (sz + 1));
        const double shx = P.L[0] * sx, shy = P.L[1] * sy, shz = P.L[2] * sz;
        for (int base = 0; base < nj; base += 64) {
            const int m = min(64, nj - base);
            double4 cx4 = make_double4(0, 0, 0, 0);
            int4 ci4 = make_int4(0, 0, 0, 0);
            int cmol = -3;
            bool near = false;
            if (l < m) {
                cx4 = xr[j0 + base + l];
                ci4 = imgs[j0 + base + l];
                cmol = mol[ci4.w];
                // lane-parallel pre-test: this j (shifted image) against the group box
                double y0 = cx4.x - P.L[0] * ci4.x + shx, y1 = cx4.y - P.L[1] * ci4.y + shy,
                       y2 = cx4.z - P.L[2] * ci4.z + shz;
                double t0 = fmax(0.0, fmax(c.lo[0] - y0, y0 - c.hi[0]));
                double t1 = fmax(0.0, fmax(c.lo[1] - y1, y1 - c.hi[1]));
                double t2 = fmax(0.0, fmax(c.lo[2] - y2, y2 - c.hi[2]));
                near = t0 * t0 + t1 * t1 + t2 * t2 <= rl2 * (1.0 + 1e-9) + 1e-9;
            }
            uint64_t todo = __ballot(near);
            while (todo) {
                const int k = __builtin_ctzll(todo);
                todo &= todo - 1;
                const int js = j0 + base + k;
                const double xjx = bcast_d(cx4.x, k), xjy = bcast_d(cx4.y, k), xjz = bcast_d(cx4.z, k);
                const int ijx = bcast_i(ci4.x, k), ijy = bcast_i(ci4.y, k), ijz = bcast_i(ci4.z, k);
                const int aj = bcast_i(ci4.w, k), molj = bcast_i(cmol, k);
                double d0 = xi.x - xjx, d1 = xi.y - xjy, d2 = xi.z - xjz;
                double n0 = anint_d(d0 * P.invL[0]), n1 = anint_d(d1 * P.invL[1]), n2 = anint_d(d2 * P.invL[2]);
                d0 = d0 - P.L[0] * n0;
                d1 = d1 - P.L[1] * n1;
                d2 = d2 - P.L[2] * n2;
                double dd = d0 * d0 + d1 * d1 + d2 * d2;
                bool in = act && (js != slot) && (dd <= P.rlim2);
                in = in && ((int)n0 - ii.x + ijx == sx) && ((int)n1 - ii.y + ijy == sy) &&
                     ((int)n2 - ii.z + ijz == sz);
                if (in && molj == moli) {
                    for (int e = xb; e < xe; ++e)
                        if (xidx[e] == aj) { in = false; ++nrem; break; }
                }
                const uint64_t mk = __ballot(in);
                if (mk) {
                    if (l == 0) {
                        Entry E;
                        E.js = (uint32_t)js | (code << JP_SLOT_BITS);
                        E.pad = 0;
                        E.m = mk;
                        tmp[pos] = E;
                    }
                    ++pos;
                    nb += in ? 1 : 0;
                }
            }
        }
    }
    if (l == 0) lcount[g * JP_WL + w] = pos - pos0;
    int t = wave_sum(nrem);
    s_nb[w][l] = nb;
    if (l == 0) s_nrem[w] = t;
    __syncthreads();
    if (w == 0) {
        int tot = 0, nbt = 0;
        for (int q = 0; q < JP_WL; ++q) { tot += s_nrem[q]; nbt += s_nb[q][l]; }
        if (l == 0 && tot) atomicAdd(&nremoved[grp_sys[g]], tot);
        if (act) nbcount[ai] = nbt;
    }
}

// tmp (per (group,wave) capacity-sized segments) -> contiguous entries.
__global__ void __launch_bounds__(256)
k_list_compact(int ngw, const int *__restrict__ coff, const int *__restrict__ lcount, const int *__restrict__ loff,
               const Entry *__restrict__ tmp, Entry *__restrict__ ent) {
    const int gw = blockIdx.x;
    if (gw >= ngw) return;
    const int n = lcount[gw], s = coff[gw], d = loff[gw];
    for (int k = threadIdx.x; k < n; k += blockDim.x) ent[d + k] = tmp[s + k];
}

// --------------------------------------------------------------------------
// per step: nonbonded pair kernels (enear_ZD_doub_PB semantics)
// --------------------------------------------------------------------------
//
// One block of WP wavefronts per i-group; lane l of every wave owns i atom
// slot 64g+l.  The block stages 64*WP list entries and their j atoms in LDS
// (coalesced entry loads, gathered 32-byte j loads, all in flight at once);
// wave w then runs over its 64 of them, every lane computing its pair with
// the j atom read by LDS broadcast.  The WP partial forces of an atom are
// combined in a fixed order at the end.

__device__ __forceinline__ void shift_of(uint32_t code, int &sx, int &sy, int &sz) {
    sx = (int)(code / 9) - 1;
    sy = (int)((code / 3) % 3) - 1;
    sz = (int)(code % 3) - 1;
}

// Plain FP64 path: md_run's per-pair arithmetic (enear.f90:182-258 ZD lines)
// in FP64 with IEEE division and square root, no FMA contraction.
template <bool ENERGY, int WP>
__global__ void __launch_bounds__(64 * WP)
k_pair_f64(int ngrp, const int *__restrict__ grp_sys, const int *__restrict__ gcnt, const SysP *__restrict__ sp,
           const JAtom *__restrict__ ys, const double *__restrict__ qd, const double2 *__restrict__ lj64,
           const int *__restrict__ loff, const Entry *__restrict__ ent, double *__restrict__ fout,
           double *__restrict__ epart) {
#pragma clang fp contract(off)
    extern __shared__ double2 ljs[];
    __shared__ Entry se[64 * WP];
    __shared__ JAtom sj[64 * WP];
    __shared__ double sq[64 * WP];
    int g = xcd_remap(blockIdx.x, gridDim.x);
    if (g >= ngrp) return;
    const SysP &P = sp[grp_sys[g]];
    const int tid = threadIdx.x, l = tid & 63, w = tid >> 6;
    for (int k = tid; k < P.ntyp * P.ntyp; k += 64 * WP) ljs[k] = lj64[P.lj0 + k];
    const int slot = g * 64 + l;
    const bool act = l < gcnt[g];
    const JAtom ai = ys[slot];
    const double qi = qd[slot];
    const int ti = ai.t & 0xfff, ci = (ai.t >> 13) & 1;
    const double f0 = P.fac[ci * 2 + 0], f1 = P.fac[ci * 2 + 1];
    const int k0 = P.cls[ci * 2 + 0], k1 = P.cls[ci * 2 + 1];
    double gx = 0.0, gy = 0.0, gz = 0.0;
    double ee[3] = {0.0, 0.0, 0.0}, ev[3] = {0.0, 0.0, 0.0};
    double xs = ai.x, ysh = ai.y, zs = ai.z;
    uint32_t cur = 13;
    const int e0 = loff[g * JP_WL], ne = loff[(g + 1) * JP_WL] - e0;
    for (int base = 0; base < ne; base += 64 * WP) {
        __syncthreads();
        if (base + tid < ne) {
            const Entry E = ent[e0 + base + tid];
            const int jsl = (int)(E.js & JP_SLOT_MASK);
            se[tid] = E;
            sj[tid] = ys[jsl];
            sq[tid] = qd[jsl];
        }
        __syncthreads();
        const int hi = min(64 * (w + 1), ne - base);
        for (int q = 64 * w; q < hi; ++q) {
            const uint32_t js = se[q].js;
            const uint64_t m = se[q].m;
            const uint32_t code = js >> JP_SLOT_BITS;
            if (code != cur) {
                int sx, sy, sz;
                shift_of(code, sx, sy, sz);
                xs = ai.x - P.L[0] * (double)sx;
                ysh = ai.y - P.L[1] * (double)sy;
                zs = ai.z - P.L[2] * (double)sz;
                cur = code;
            }
            const JAtom aj = sj[q];
            const double qj = sq[q];
            const int tj = aj.t & 0xfff, cj = (aj.t >> 13) & 1;
            const bool vdwj = (aj.t >> 12) & 1;
            double dx = xs - aj.x, dy = ysh - aj.y, dz = zs - aj.z;
            if (P.small) {
                // edge < 2*rlim: the image chosen at the rebuild may stop being
                // the minimum one before the next rebuild (cutoff.f90 tests anint
                // images; enear.f90:189-191 re-takes them every step)
                dx = dx - P.L[0] * anint_d(dx * P.invL[0]);
                dy = dy - P.L[1] * anint_d(dy * P.invL[1]);
                dz = dz - P.L[2] * anint_d(dz * P.invL[2]);
            }
            double dd2 = dx * dx + dy * dy + dz * dz;
            const bool in = (m >> l) & 1;
            dd2 = in ? dd2 : 1.0;
            double dd2i = 1.0 / dd2;
            double co = (dd2i - P.irlim2 >= 0.0) ? 1.0 : 0.0;
            co = in ? co : 0.0;
            double ddi = sqrt(dd2i);
            double eel = qi * qj * co;
            double cc = eel * (dd2i * ddi - P.fcoeff);
            double vv = 0.0;
            if (vdwj) {
                const double2 c = ljs[ti * P.ntyp + tj];
                double dd6i = dd2i * dd2i * dd2i * co;
                double v1 = c.x * dd6i;
                double v2 = c.y * dd6i * dd6i;
                vv = v2 - v1;
                cc = cc + (-6.0 * v1 + 12.0 * v2) * dd2i;
            }
            const double fc = cj ? f1 : f0;
            cc = cc * fc;
            gx = gx + dx * cc;
            gy = gy + dy * cc;
            gz = gz + dz * cc;
            if (ENERGY) {
                double el = eel * (ddi + P.bcoeff * dd2 - P.zcore);
                const int k = cj ? k1 : k0;
                ee[0] += (k == 0) ? el : 0.0; ee[1] += (k == 1) ? el : 0.0; ee[2] += (k == 2) ? el : 0.0;
                ev[0] += (k == 0) ? vv : 0.0; ev[1] += (k == 1) ? vv : 0.0; ev[2] += (k == 2) ? vv : 0.0;
            }
        }
    }
    // combine the WP partial sums of each atom in wave order
    __syncthreads();
    double *red = (double *)sj;   // 64*WP*32 B >= WP*64*3*8 B
    red[(w * 64 + l) * 3] = gx;
    red[(w * 64 + l) * 3 + 1] = gy;
    red[(w * 64 + l) * 3 + 2] = gz;
    __syncthreads();
    if (w == 0 && act) {
        double sx = 0.0, sy = 0.0, sz = 0.0;
        for (int q = 0; q < WP; ++q) {
            sx = sx + red[(q * 64 + l) * 3];
            sy = sy + red[(q * 64 + l) * 3 + 1];
            sz = sz + red[(q * 64 + l) * 3 + 2];
        }
        fout[3 * (size_t)slot] = 0.0 - sx;
        fout[3 * (size_t)slot + 1] = 0.0 - sy;
        fout[3 * (size_t)slot + 2] = 0.0 - sz;
    }
    if (ENERGY) {
        __syncthreads();
        double *re = (double *)se;   // 64*WP*16 B >= WP*6*8 B
        for (int k = 0; k < 3; ++k) {
            double a = wave_sum(act ? ee[k] : 0.0), b = wave_sum(act ? ev[k] : 0.0);
            if (l == 0) { re[w * 6 + k] = a; re[w * 6 + 3 + k] = b; }
        }
        __syncthreads();
        if (tid < 6) {
            double t = 0.0;
            for (int q = 0; q < WP; ++q) t = t + re[q * 6 + tid];
            epart[6 * (size_t)g + tid] = t;
        }
    }
}

// FP32 path: coordinate differences in FP64 (y is up to ~200 A), everything
// after them in FP32; per-lane FP32 partial sums are flushed every JP_FLUSH j
// into 64-bit fixed point (deterministic, independent of summation order).
template <bool ENERGY, int WP>
__global__ void __launch_bounds__(64 * WP)
k_pair_f32(int ngrp, const int *__restrict__ grp_sys, const int *__restrict__ gcnt, const SysP *__restrict__ sp,
           const JAtom *__restrict__ ys, const float2 *__restrict__ lj32, const int *__restrict__ loff,
           const Entry *__restrict__ ent, long long *__restrict__ fout, long long *__restrict__ epart) {
    extern __shared__ float2 ljf[];
    __shared__ Entry se[64 * WP];
    __shared__ JAtom sj[64 * WP];
    int g = xcd_remap(blockIdx.x, gridDim.x);
    if (g >= ngrp) return;
    const SysP &P = sp[grp_sys[g]];
    const int tid = threadIdx.x, l = tid & 63, w = tid >> 6;
    for (int k = tid; k < P.ntyp * P.ntyp; k += 64 * WP) ljf[k] = lj32[P.lj0 + k];
    const int slot = g * 64 + l;
    const bool act = l < gcnt[g];
    const JAtom ai = ys[slot];
    const float qi = ai.q;
    const int ti = ai.t & 0xfff, ci = (ai.t >> 13) & 1;
    const float f0 = (float)P.fac[ci * 2 + 0], f1 = (float)P.fac[ci * 2 + 1];
    const int k0 = P.cls[ci * 2 + 0], k1 = P.cls[ci * 2 + 1];
    const float rc2 = (float)P.rc2, fcf = (float)P.fcoeff, bcf = (float)P.bcoeff, zcf = (float)P.zcore;
    float fx = 0.f, fy = 0.f, fz = 0.f;
    long long ax = 0, ay = 0, az = 0;
    float ee[3] = {0.f, 0.f, 0.f}, ev[3] = {0.f, 0.f, 0.f};
    long long eacc[6] = {0, 0, 0, 0, 0, 0};
    double xs = ai.x, ysh = ai.y, zs = ai.z;
    uint32_t cur = 13;
    int nf = 0;
    const int e0 = loff[g * JP_WL], ne = loff[(g + 1) * JP_WL] - e0;
    for (int base = 0; base < ne; base += 64 * WP) {
        __syncthreads();
        if (base + tid < ne) {
            const Entry E = ent[e0 + base + tid];
            se[tid] = E;
            sj[tid] = ys[E.js & JP_SLOT_MASK];
        }
        __syncthreads();
        const float *ljrow = (const float *)(ljf + ti * P.ntyp);
        const int hi = min(64 * (w + 1), ne - base);
        for (int q = 64 * w; q < hi; ++q) {
            const uint32_t js = se[q].js;
            const uint64_t m = se[q].m;
            const uint32_t code = js >> JP_SLOT_BITS;
            if (code != cur) {
                int sx, sy, sz;
                shift_of(code, sx, sy, sz);
                xs = ai.x - P.L[0] * (double)sx;
                ysh = ai.y - P.L[1] * (double)sy;
                zs = ai.z - P.L[2] * (double)sz;
                cur = code;
            }
            const JAtom aj = sj[q];
            const int tj = aj.t & 0xfff, cj = (aj.t >> 13) & 1;
            double ddx = xs - aj.x, ddy = ysh - aj.y, ddz = zs - aj.z;
            if (P.small) {
                ddx = ddx - P.L[0] * rint(ddx * P.invL[0]);
                ddy = ddy - P.L[1] * rint(ddy * P.invL[1]);
                ddz = ddz - P.L[2] * rint(ddz * P.invL[2]);
            }
            const float dx = (float)ddx, dy = (float)ddy, dz = (float)ddz;
            float r2 = dx * dx + dy * dy + dz * dz;
            const bool in = ((m >> l) & 1) && (r2 <= rc2);
            r2 = in ? r2 : 1.0f;
            const float ri = __builtin_amdgcn_rsqf(r2);
            const float ri2 = ri * ri;
            const float qq = qi * aj.q;
            float c = qq * (ri * ri2 - fcf);
            float vv = 0.f;
            if ((aj.t >> 12) & 1) {
                const float c6 = ljrow[2 * tj], c12 = ljrow[2 * tj + 1];
                const float ri6 = ri2 * ri2 * ri2;
                const float v1 = c6 * ri6, v2 = c12 * ri6 * ri6;
                c += (12.f * v2 - 6.f * v1) * ri2;
                vv = v2 - v1;
            }
            c = in ? c * (cj ? f1 : f0) : 0.f;
            fx += c * dx;
            fy += c * dy;
            fz += c * dz;
            if (ENERGY) {
                float el = in ? qq * (ri + bcf * r2 - zcf) : 0.f;
                vv = in ? vv : 0.f;
                const int k = cj ? k1 : k0;
                ee[0] += (k == 0) ? el : 0.f; ee[1] += (k == 1) ? el : 0.f; ee[2] += (k == 2) ? el : 0.f;
                ev[0] += (k == 0) ? vv : 0.f; ev[1] += (k == 1) ? vv : 0.f; ev[2] += (k == 2) ? vv : 0.f;
            }
            if (++nf == JP_FLUSH) {
                nf = 0;
                ax += f2fix(fx, JP_FSCALE); ay += f2fix(fy, JP_FSCALE); az += f2fix(fz, JP_FSCALE);
                fx = fy = fz = 0.f;
                if (ENERGY)
                    for (int k = 0; k < 3; ++k) {
                        eacc[k] += f2fix(ee[k], JP_ESCALE); eacc[3 + k] += f2fix(ev[k], JP_ESCALE);
                        ee[k] = ev[k] = 0.f;
                    }
            }
        }
    }
    ax += f2fix(fx, JP_FSCALE); ay += f2fix(fy, JP_FSCALE); az += f2fix(fz, JP_FSCALE);
    __syncthreads();
    long long *red = (long long *)sj;   // 64*WP*32 B >= WP*64*3*8 B
    red[(w * 64 + l) * 3] = ax;
    red[(w * 64 + l) * 3 + 1] = ay;
    red[(w * 64 + l) * 3 + 2] = az;
    __syncthreads();
    if (w == 0 && act) {
        long long sx = 0, sy = 0, sz = 0;
        for (int q = 0; q < WP; ++q) {
            sx += red[(q * 64 + l) * 3];
            sy += red[(q * 64 + l) * 3 + 1];
            sz += red[(q * 64 + l) * 3 + 2];
        }
        fout[3 * (size_t)slot] = -sx;
        fout[3 * (size_t)slot + 1] = -sy;
        fout[3 * (size_t)slot + 2] = -sz;
    }
    if (ENERGY) {
        for (int k = 0; k < 3; ++k) {
            eacc[k] += f2fix(ee[k], JP_ESCALE);
            eacc[3 + k] += f2fix(ev[k], JP_ESCALE);
        }
        __syncthreads();
        long long *re = (long long *)se;
        for (int k = 0; k < 6; ++k) {
            long long t = wave_sum(act ? eacc[k] : 0ll);
            if (l == 0) re[w * 6 + k] = t;
        }
        __syncthreads();
        if (tid < 6) {
            long long t = 0;
            for (int q = 0; q < WP; ++q) t += re[q * 6 + tid];
            epart[6 * (size_t)g + tid] = t;
        }
    }
}

// --------------------------------------------------------------------------
// per step: integration (verlet.f90:144-215) with the ZD excess term
// --------------------------------------------------------------------------

// Excess (excluded-pair) gradient of atom a (enear.f90:528-550), FP64, in
// md_run's accumulation order, scaled per energy class.
__device__ __forceinline__ void excess_grad(int a, const SysP &P, const double *__restrict__ x,
                                            const double *__restrict__ chg, const int *__restrict__ eptr,
                                            const int2 *__restrict__ eidx, const double *__restrict__ efac,
                                            double &g0, double &g1, double &g2) {
#pragma clang fp contract(off)
    for (int e = eptr[a]; e < eptr[a + 1]; ++e) {
        const int2 pe = eidx[e];   // partner, class
        const int b = pe.x;
        double dx = x[3 * a] - x[3 * b], dy = x[3 * a + 1] - x[3 * b + 1], dz = x[3 * a + 2] - x[3 * b + 2];
        dx = dx - P.L[0] * anint_d(dx * P.invL[0]);
        dy = dy - P.L[1] * anint_d(dy * P.invL[1]);
        dz = dz - P.L[2] * anint_d(dz * P.invL[2]);
        double cc = chg[a] * chg[b];
        cc = cc * P.fcoeff;
        double s = efac[pe.y];
        g0 = g0 + dx * cc * s;
        g1 = g1 + dy * cc * s;
        g2 = g2 + dz * cc * s;
    }
}

// gradient = excess + pair (sorted slot), then kick and drift.  With
// `gout` the gradient is also stored (original order) for export.
template <int PREC>
__global__ void k_kick(int natom, const int *__restrict__ atom_sys, const SysP *__restrict__ sp,
                       const int *__restrict__ inum, const double *__restrict__ f64, const long long *__restrict__ ffix,
                       const double *__restrict__ chg, const int *__restrict__ eptr, const int2 *__restrict__ eidx,
                       const double *__restrict__ efac_all, const double *__restrict__ mass, double *__restrict__ x,
                       double *__restrict__ v, double *__restrict__ xold, double *__restrict__ vback,
                       double *__restrict__ gout, int do_step) {
#pragma clang fp contract(off)
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natom) return;
    const int s = atom_sys[a];
    const SysP &P = sp[s];
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    excess_grad(a, P, x, chg, eptr, eidx, efac_all + 3 * s, g0, g1, g2);
    const int sl = inum[a];
    if (PREC == 0) {
        g0 = g0 + f64[3 * (size_t)sl];
        g1 = g1 + f64[3 * (size_t)sl + 1];
        g2 = g2 + f64[3 * (size_t)sl + 2];
    } else {
        g0 = g0 + (double)ffix[3 * (size_t)sl] * (1.0 / JP_FSCALE);
        g1 = g1 + (double)ffix[3 * (size_t)sl + 1] * (1.0 / JP_FSCALE);
        g2 = g2 + (double)ffix[3 * (size_t)sl + 2] * (1.0 / JP_FSCALE);
    }
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
__global__ void k_shake2(int ng, const int *__restrict__ gsys, const SysP *__restrict__ sp,
                         const int4 *__restrict__ ga, const double *__restrict__ gd, const double *__restrict__ mass,
                         double *__restrict__ r1, const double *__restrict__ r2) {
#pragma clang fp contract(off)
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    const SysP &P = sp[gsys[g]];
    const int ia = ga[g].x, ja = ga[g].y;
    const double im1 = 1.0 / mass[ia], im2 = 1.0 / mass[ja];
    const double im3 = im1 + im2, im4 = im3 * im3;
    double d1[3], d2[3];
    for (int k = 0; k < 3; ++k) {
        d1[k] = r1[3 * ja + k] - r1[3 * ia + k];
        d2[k] = r2[3 * ja + k] - r2[3 * ia + k];
    }
    mimg3(P, d1);
    mimg3(P, d2);
    double d1d1 = d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2];
    double d1d2 = d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2];
    double d2d2 = d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2];
    double A = im4 * d2d2, B = im3 * d1d2, C = d1d1 - gd[6 * g];
    double G = (B - sqrt(B * B - C * A)) / A;
    for (int k = 0; k < 3; ++k) {
        r1[3 * ia + k] = r1[3 * ia + k] + (im1 * G * d2[k]);
        r1[3 * ja + k] = r1[3 * ja + k] - (im2 * G * d2[k]);
    }
}

// shake3_PB, shake.f90:112-327 (3-atom groups, Newton with exact 3x3 inverse)
__global__ void k_shake3(int ng, const int *__restrict__ gsys, const SysP *__restrict__ sp,
                         const int4 *__restrict__ ga, const double *__restrict__ gd,
                         const double *__restrict__ imass, double *__restrict__ r1, const double *__restrict__ r2,
                         double tol, int maxit, int *__restrict__ err) {
#pragma clang fp contract(off)
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    const SysP &P = sp[gsys[g]];
    const int a[3] = {ga[g].x, ga[g].y, ga[g].z};
    const double im[3] = {imass[a[0]], imass[a[1]], imass[a[2]]};
    double co[3][3];
    co[0][0] = im[0] + im[1]; co[0][1] = -im[1]; co[0][2] = -im[0];
    co[1][0] = -im[1]; co[1][1] = im[1] + im[2]; co[1][2] = -im[2];
    co[2][0] = -im[0]; co[2][1] = -im[2]; co[2][2] = im[2] + im[0];
    const double T[3] = {gd[6 * g], gd[6 * g + 1], gd[6 * g + 2]};
    const double TT[3] = {tol * T[0], tol * T[1], tol * T[2]};
    double G[6] = {0, 0, 0, 0, 0, 0};
    double d1[3][3], d2[3][3];
    for (int k = 0; k < 3; ++k) {
        d1[0][k] = r1[3 * a[0] + k] - r1[3 * a[1] + k];
        d1[1][k] = r1[3 * a[1] + k] - r1[3 * a[2] + k];
        d1[2][k] = r1[3 * a[2] + k] - r1[3 * a[0] + k];
        d2[0][k] = r2[3 * a[0] + k] - r2[3 * a[1] + k];
        d2[1][k] = r2[3 * a[1] + k] - r2[3 * a[2] + k];
        d2[2][k] = r2[3 * a[2] + k] - r2[3 * a[0] + k];
    }
    for (int q = 0; q < 3; ++q) { mimg3(P, d1[q]); mimg3(P, d2[q]); }
    for (int it = 0; it < maxit; ++it) {
        double B[3];
        for (int q = 0; q < 3; ++q) B[q] = T[q] - (d1[q][0] * d1[q][0] + d1[q][1] * d1[q][1] + d1[q][2] * d1[q][2]);
        if (fabs(B[0]) < TT[0] && fabs(B[1]) < TT[1] && fabs(B[2]) < TT[2]) {
            for (int k = 0; k < 3; ++k) {
                r1[3 * a[0] + k] = r1[3 * a[0] + k] + im[0] * (G[3] * d2[0][k] - G[5] * d2[2][k]);
                r1[3 * a[1] + k] = r1[3 * a[1] + k] + im[1] * (-G[3] * d2[0][k] + G[4] * d2[1][k]);
                r1[3 * a[2] + k] = r1[3 * a[2] + k] + im[2] * (-G[4] * d2[1][k] + G[5] * d2[2][k]);
            }
            return;
        }
        double gf[3][3];
        for (int ig = 0; ig < 3; ++ig)
            for (int q = 0; q < 3; ++q)
                gf[q][ig] = 2.0 * co[q][ig] * (d1[q][0] * d2[ig][0] + d1[q][1] * d2[ig][1] + d1[q][2] * d2[ig][2]);
        double deta = (gf[0][0] * gf[1][1] * gf[2][2]) + (gf[0][1] * gf[1][2] * gf[2][0]) +
                      (gf[0][2] * gf[1][0] * gf[2][1]) - (gf[0][0] * gf[1][2] * gf[2][1]) -
                      (gf[0][1] * gf[1][0] * gf[2][2]) - (gf[0][2] * gf[1][1] * gf[2][0]);
        if (deta == 0.0) { atomicOr(err, 8); return; }
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
    atomicOr(err, 4);  // not converged
}

// v = (x - xold)/dt after SHAKE (verlet.f90:203-211) and the per-step refresh
// of the sorted pair-kernel coordinates (enear.f90:72-76 re-sorts every call).
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

// Unique pairs with r <= rc (each pair once) at the current coordinates:
// out[0] all pairs, out[1] pairs where both atoms carry LJ, out[2] listed
// pairs (r <= rlim at the last rebuild, each once).
__global__ void __launch_bounds__(64)
k_count_pairs(int ngrp, const int *__restrict__ grp_sys, const int *__restrict__ gcnt, const SysP *__restrict__ sp,
              const JAtom *__restrict__ ys, const int *__restrict__ loff,
              const Entry *__restrict__ ent, unsigned long long *__restrict__ out) {
    int g = blockIdx.x;
    if (g >= ngrp) return;
    const SysP &P = sp[grp_sys[g]];
    const int l = threadIdx.x, slot = g * 64 + l;
    const JAtom ai = ys[slot];
    const Entry *E = ent + loff[g * JP_WL];
    const int ne = loff[(g + 1) * JP_WL] - loff[g * JP_WL];
    unsigned long long c = 0, clj = 0, cl = 0;
    for (int e = 0; e < ne; ++e) {
        const int jsl = (int)(E[e].js & JP_SLOT_MASK);
        int sx, sy, sz;
        shift_of(E[e].js >> JP_SLOT_BITS, sx, sy, sz);
        const JAtom aj = ys[jsl];
        double dx = ai.x - P.L[0] * sx - aj.x, dy = ai.y - P.L[1] * sy - aj.y, dz = ai.z - P.L[2] * sz - aj.z;
        if (P.small) {
            dx = dx - P.L[0] * rint(dx * P.invL[0]);
            dy = dy - P.L[1] * rint(dy * P.invL[1]);
            dz = dz - P.L[2] * rint(dz * P.invL[2]);
        }
        bool lis = ((E[e].m >> l) & 1) && (slot < jsl);
        bool in = lis && (dx * dx + dy * dy + dz * dz <= P.rc2);
        c += in;
        clj += in && ((ai.t >> 12) & 1) && ((aj.t >> 12) & 1);
        cl += lis;
    }
    c = wave_sum(c);
    clj = wave_sum(clj);
    cl = wave_sum(cl);
    if (l == 0) {
        atomicAdd(out, c);
        atomicAdd(out + 1, clj);
        atomicAdd(out + 2, cl);
    }
}
