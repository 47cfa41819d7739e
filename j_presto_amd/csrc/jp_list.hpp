// Pair-list rebuild on 8-atom i-clusters (cutoff.f90 cutatm_PB + cutoff_PB
// semantics).
//
// The list of i-cluster C (slots 8C..8C+7) is a run of entries
//   uint2 { j slot | image code << 27,  8-bit mask }
// bit k of the mask set iff (slot 8C+k, j) is in md_run's pair list: the FP64
// minimum-image distance on raw coordinates is <= rlim (cutoff.f90:464-469),
// j != i, and (i,j) is not an excluded pair (cutoff.f90:50-96).  The image
// code gives the periodic shift s (code = (sx+1)*9+(sy+1)*3+(sz+1)); the
// pair kernels use j at y_j + s*L.
//
// Entries whose j carries LJ are stored from the front of the cluster's
// segment [coff[C], coff[C+1]), the others from its back, so that every chunk
// the pair kernels stage is uniformly with or without the LJ term.
//
// Candidates: every j-cluster whose home cell (the cell of its first atom) is
// within rlim + 2*ext of C's box, ext bounding how far a cluster's atoms
// stick out of its home cell, then every j-cluster whose own box (shifted by
// s*L) is within rlim of C's box.  Pipeline: k_cbox -> k_list_cap8 (count
// candidate j-clusters) -> scan -> k_list8 (one wavefront per i-cluster; the
// 8 i atoms sit in scalar registers, each lane tests one candidate j atom).
//
// The per-pair test runs in FP32 on coordinates relative to C's box centre.
// Its error is bounded by 2^-24 * (|r_i| + |r_j| + |d|) per component, i.e.
// below 3e-4 A^2 on d^2 for any cluster smaller than 100 A; pairs with
// |d^2 - rlim^2| <= JP_BAND (0.01 A^2) are decided by md_run's FP64 test on
// raw coordinates, so the list is the reference's bit for bit.  In boxes with
// an edge below 2*rlim every pair takes the FP64 test, which also checks that
// s is the minimum image.
#pragma once
#include "jp_kernels.hpp"

#define JP_CL 8          // atoms per i-cluster
#define JP_LW 4          // i-clusters (wavefronts) per block in the list kernels
#define JP_QCAP 128      // per-wave queue of candidate j-clusters (power of 2)
#define JP_BAND 0.01     // A^2: FP32 results this close to rlim^2 go to the FP64 test
#define JP_ERR_LISTCAP 16

// Orders LDS accesses of one wavefront (lanes write, other lanes read).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int lane_rank(uint64_t b) {   // set bits of b below this lane
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u));
}

__host__ __device__ __forceinline__ void shift_of(uint32_t code, int &sx, int &sy, int &sz) {
    sx = (int)(code / 9) - 1;
    sy = (int)((code / 3) % 3) - 1;
    sz = (int)(code % 3) - 1;
}

// Box and centre of every cluster (wrapped coordinates at this rebuild), and
// how far it sticks out of its home cell: maximum per cell and per system
// (non-negative floats compare like their bit patterns).
__global__ void k_cbox(int nclu, const int *__restrict__ clu_sys, const SysP *__restrict__ sp,
                       const int *__restrict__ perm, const int *__restrict__ cellid, const int *__restrict__ ccoord,
                       const JAtom *__restrict__ ys, double *__restrict__ cbb, double4 *__restrict__ ccen,
                       unsigned *__restrict__ cellext, unsigned *__restrict__ sysext) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nclu) return;
    const int s = clu_sys[c];
    const SysP &P = sp[s];
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    int n = 0;
    for (int k = 0; k < JP_CL; ++k) {
        if (perm[c * JP_CL + k] < 0) continue;
        const JAtom a = ys[c * JP_CL + k];
        lo[0] = fmin(lo[0], a.x); hi[0] = fmax(hi[0], a.x);
        lo[1] = fmin(lo[1], a.y); hi[1] = fmax(hi[1], a.y);
        lo[2] = fmin(lo[2], a.z); hi[2] = fmax(hi[2], a.z);
        ++n;
    }
    for (int k = 0; k < 3; ++k) { cbb[6 * c + k] = lo[k]; cbb[6 * c + 3 + k] = hi[k]; }
    if (n == 0) { ccen[c] = make_double4(0.0, 0.0, 0.0, 0.0); return; }
    ccen[c] = make_double4(0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1]), 0.5 * (lo[2] + hi[2]), 0.0);
    // slots are real atoms first, padding last: a non-empty cluster's first slot is real
    const int cell = cellid[perm[c * JP_CL]];
    const int cc = ccoord[cell];
    const int cx[3] = {cc & 1023, (cc >> 10) & 1023, (cc >> 20) & 1023};
    double e = 0.0;
    for (int k = 0; k < 3; ++k) {
        const double clo = P.wall[k] + cx[k] * P.csz[k], chi = P.wall[k] + (cx[k] + 1) * P.csz[k];
        e = fmax(e, fmax(clo - lo[k], hi[k] - chi));
    }
    const float ef = (float)e * 1.0001f + 1e-5f;
    atomicMax(cellext + cell, __float_as_uint(ef));
    atomicMax(sysext + s, __float_as_uint(ef));
}

struct CluRange {
    int ulo[3], n[3], ncand;
    double lo[3], hi[3], rl;
    bool bad;
};

// Unwrapped cell range around cluster ci's box grown by rlim + 2*ext.
__device__ __forceinline__ CluRange clu_range(const SysP &P, const double *__restrict__ cbb, int ci, float sysext) {
    CluRange r;
    r.rl = sqrt(P.rlim2) * (1.0 + 1e-12) + 1e-9;   // conservative pruning radius
    r.bad = false;
    bool empty = false;
    const double m = r.rl + 2.0 * (double)sysext;
    for (int k = 0; k < 3; ++k) {
        r.lo[k] = cbb[6 * ci + k];
        r.hi[k] = cbb[6 * ci + 3 + k];
        if (!(r.lo[k] <= r.hi[k])) empty = true;
        const double a = floor((r.lo[k] - m - P.wall[k]) * P.icsz[k]);
        const double b = floor((r.hi[k] + m - P.wall[k]) * P.icsz[k]);
        // non-finite or runaway coordinates: report and do nothing (bounded loops)
        if (!empty && (!(a >= -4.0 * P.nc[k] - 64.0) || !(b <= 5.0 * P.nc[k] + 64.0) || !(b - a <= 3.0 * P.nc[k] + 16.0)))
            r.bad = true;
        r.ulo[k] = (r.bad || empty) ? 0 : (int)a;
        r.n[k] = (r.bad || empty) ? 0 : (int)(b - a) + 1;
    }
    r.ncand = (r.bad || empty) ? 0 : r.n[0] * r.n[1] * r.n[2];
    return r;
}

__device__ __forceinline__ int floordiv(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

// Candidate q -> global cell id and image shift, or -1 when the unwrapped cell
// (grown by 2*its ext) is farther than rlim from the cluster box.  Shifts of
// two or more boxes are never within rlim (every edge is >= 2*cutoff).
__device__ __forceinline__ int clu_cell(const SysP &P, const CluRange &r, const int *__restrict__ crank,
                                        const unsigned *__restrict__ cellext, int q, uint32_t &code, double &shx,
                                        double &shy, double &shz) {
    const int ux = r.ulo[0] + q % r.n[0], uy = r.ulo[1] + (q / r.n[0]) % r.n[1], uz = r.ulo[2] + q / (r.n[0] * r.n[1]);
    const int sx = floordiv(ux, P.nc[0]), sy = floordiv(uy, P.nc[1]), sz = floordiv(uz, P.nc[2]);
    if (sx < -1 || sx > 1 || sy < -1 || sy > 1 || sz < -1 || sz > 1) return -1;
    const int cx = ux - sx * P.nc[0], cy = uy - sy * P.nc[1], cz = uz - sz * P.nc[2];
    const int cell = crank[P.cell0 + (cz * P.nc[1] + cy) * P.nc[0] + cx];
    const double m = r.rl + 2.0 * (double)__uint_as_float(cellext[cell]);
    const double gx = fmax(0.0, fmax(P.wall[0] + ux * P.csz[0] - r.hi[0], r.lo[0] - (P.wall[0] + (ux + 1) * P.csz[0])));
    const double gy = fmax(0.0, fmax(P.wall[1] + uy * P.csz[1] - r.hi[1], r.lo[1] - (P.wall[1] + (uy + 1) * P.csz[1])));
    const double gz = fmax(0.0, fmax(P.wall[2] + uz * P.csz[2] - r.hi[2], r.lo[2] - (P.wall[2] + (uz + 1) * P.csz[2])));
    if (gx * gx + gy * gy + gz * gz > m * m) return -1;
    code = (uint32_t)((sx + 1) * 9 + (sy + 1) * 3 + (sz + 1));
    shx = P.L[0] * sx;
    shy = P.L[1] * sy;
    shz = P.L[2] * sz;
    return cell;
}

// Box of j-cluster jc, shifted by (shx,shy,shz), within rl of the i-cluster box?
__device__ __forceinline__ bool clu_near(const double *__restrict__ cbb, int jc, double shx, double shy, double shz,
                                         const CluRange &r) {
    const double *b = cbb + 6 * (size_t)jc;
    const double gx = fmax(0.0, fmax(b[0] + shx - r.hi[0], r.lo[0] - (b[3] + shx)));
    const double gy = fmax(0.0, fmax(b[1] + shy - r.hi[1], r.lo[1] - (b[4] + shy)));
    const double gz = fmax(0.0, fmax(b[2] + shz - r.hi[2], r.lo[2] - (b[5] + shz)));
    return gx * gx + gy * gy + gz * gz <= r.rl * r.rl;
}

// j-clusters whose first slot lies in `cell`
__device__ __forceinline__ void cell_clusters(const int *__restrict__ cslot, const int *__restrict__ ccnt, int cell,
                                              int &k0, int &k1) {
    const int a = cslot[cell], n = ccnt[cell];
    k0 = (a + JP_CL - 1) / JP_CL;
    k1 = (a + n + JP_CL - 1) / JP_CL;
}

// md_run's list test on raw coordinates (cutoff.f90:464-469) plus: s is the
// pair's minimum image.
__device__ __forceinline__ bool exact_in(const SysP &P, const double4 &a, const int4 &ia, const double4 &b,
                                         const int4 &ib, int sx, int sy, int sz) {
#pragma clang fp contract(off)
    double d0 = a.x - b.x, d1 = a.y - b.y, d2 = a.z - b.z;
    const double n0 = anint_d(d0 * P.invL[0]), n1 = anint_d(d1 * P.invL[1]), n2 = anint_d(d2 * P.invL[2]);
    d0 = d0 - P.L[0] * n0;
    d1 = d1 - P.L[1] * n1;
    d2 = d2 - P.L[2] * n2;
    const double dd = d0 * d0 + d1 * d1 + d2 * d2;
    return dd <= P.rlim2 && ((int)n0 - ia.x + ib.x == sx) && ((int)n1 - ia.y + ib.y == sy) &&
           ((int)n2 - ia.z + ib.z == sz);
}

// Upper bound of each i-cluster's entries: 8 per candidate j-cluster.
__global__ void __launch_bounds__(64 * JP_LW)
k_list_cap8(int nclu, const int *__restrict__ clu_sys, const SysP *__restrict__ sp, const double *__restrict__ cbb,
            const int *__restrict__ crank, const int *__restrict__ cslot, const int *__restrict__ ccnt,
            const unsigned *__restrict__ cellext, const unsigned *__restrict__ sysext, int *__restrict__ cap,
            int *__restrict__ err) {
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int ci = blockIdx.x * JP_LW + w;
    if (ci >= nclu) return;
    const int s = clu_sys[ci];
    const SysP &P = sp[s];
    const CluRange r = clu_range(P, cbb, ci, __uint_as_float(sysext[s]));
    if (r.bad && l == 0) atomicOr(err, 1);
    int cnt = 0;
    for (int qb = 0; qb < r.ncand; qb += 64) {
        const int q = qb + l;
        if (q >= r.ncand) continue;
        uint32_t code = 13;
        double shx = 0.0, shy = 0.0, shz = 0.0;
        const int cell = clu_cell(P, r, crank, cellext, q, code, shx, shy, shz);
        if (cell < 0) continue;
        int k0, k1;
        cell_clusters(cslot, ccnt, cell, k0, k1);
        for (int jc = k0; jc < k1; ++jc) cnt += clu_near(cbb, jc, shx, shy, shz, r) ? 1 : 0;
    }
    cnt = wave_sum(cnt);
    if (l == 0) cap[ci] = cnt * JP_CL;
}

// One wavefront per i-cluster.  Candidate j-clusters go through a queue in
// LDS; every 8 of them are tested as one batch (lane = one j atom against the
// 8 i atoms).  Per i atom the listed count (nbcount) and per system the number
// of excluded pairs met within rlim (nremoved) are reported for md_run's
// checks.
__global__ void __launch_bounds__(64 * JP_LW)
k_list8(int nclu, const int *__restrict__ clu_sys, const SysP *__restrict__ sp, const double *__restrict__ cbb,
        const double4 *__restrict__ ccen, const int *__restrict__ crank, const int *__restrict__ cslot,
        const int *__restrict__ ccnt, const unsigned *__restrict__ cellext, const unsigned *__restrict__ sysext,
        const JAtom *__restrict__ ys, const double4 *__restrict__ xr, const int4 *__restrict__ imgs,
        const int *__restrict__ smol, const int *__restrict__ xptr, const int *__restrict__ xidx,
        const int *__restrict__ coff, uint2 *__restrict__ ent, int2 *__restrict__ lcnt, int *__restrict__ nremoved,
        int *__restrict__ nbcount, int *__restrict__ err) {
    __shared__ uint32_t s_q[JP_LW][JP_QCAP];
    __shared__ double4 s_xr[JP_LW][JP_CL];
    __shared__ int4 s_im[JP_LW][JP_CL];
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int ci = blockIdx.x * JP_LW + w;
    if (ci >= nclu) return;   // whole wavefront; no block-wide barrier below
    const int s = clu_sys[ci];
    const SysP &P = sp[s];
    const CluRange r = clu_range(P, cbb, ci, __uint_as_float(sysext[s]));
    const double4 cen = ccen[ci];

    // the 8 i atoms: FP32 relative coordinates, molecule, atom index
    float fx = 1e15f, fy = 1e15f, fz = 1e15f;
    int mi = -2, ai = -1;
    if (l < JP_CL) {
        const int sl = ci * JP_CL + l;
        const int4 im = imgs[sl];
        ai = im.w;
        if (ai >= 0) {
            const JAtom a = ys[sl];
            fx = (float)(a.x - cen.x);
            fy = (float)(a.y - cen.y);
            fz = (float)(a.z - cen.z);
            mi = smol[sl];
        }
        s_xr[w][l] = xr[sl];
        s_im[w][l] = im;
    }
    // uniform copies (scalar registers); lane k keeps atom k's index and count
    float xi[JP_CL], yi[JP_CL], zi[JP_CL];
    int moli[JP_CL];
#pragma unroll
    for (int k = 0; k < JP_CL; ++k) {
        xi[k] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(fx), k));
        yi[k] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(fy), k));
        zi[k] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(fz), k));
        moli[k] = __builtin_amdgcn_readlane(mi, k);
    }
    const uint32_t ivalid = (uint32_t)__ballot(l < JP_CL && ai >= 0);
    int nbv = 0;
    wave_lds_sync();

    const int e0 = coff[ci], cap = coff[ci + 1] - e0;
    const float rlo = (float)(P.rlim2 - JP_BAND), rhi = (float)(P.rlim2 + JP_BAND);
    const bool small = P.small != 0;
    int nA = 0, nB = 0, nrem = 0, qh = 0, qt = 0;
    bool ovf = false;
    const int jsub = l >> 3, ja = l & 7;

    // test j-clusters s_q[qh .. qh+nq) (nq <= 8): lane = atom ja of entry jsub
    auto drain = [&](int nq) {
        const bool qv = jsub < nq;
        const uint32_t qe = qv ? s_q[w][(qh + jsub) & (JP_QCAP - 1)] : 0u;
        const int jc = (int)(qe & JP_SLOT_MASK);
        const uint32_t code = qe >> JP_SLOT_BITS;
        const int js = jc * JP_CL + ja;
        int sx, sy, sz;
        shift_of(code, sx, sy, sz);
        uint32_t m = 0, amb = 0;
        int aj = -1, molj = -3, tj = 0;
        int4 ij = make_int4(0, 0, 0, -1);
        if (qv) {
            ij = imgs[js];
            aj = ij.w;
        }
        if (aj >= 0) {
            const JAtom b = ys[js];
            tj = b.t;
            molj = smol[js];
            const float rx = (float)((b.x + P.L[0] * sx) - cen.x);
            const float ry = (float)((b.y + P.L[1] * sy) - cen.y);
            const float rz = (float)((b.z + P.L[2] * sz) - cen.z);
#pragma unroll
            for (int k = 0; k < JP_CL; ++k) {
                const float dx = xi[k] - rx, dy = yi[k] - ry, dz = zi[k] - rz;
                const float d2 = dx * dx + dy * dy + dz * dz;
                m |= (d2 < rlo ? 1u : 0u) << k;
                amb |= ((d2 >= rlo) && (d2 <= rhi) ? 1u : 0u) << k;
            }
            if (small) { m = 0; amb = 0xffu; }
            m &= ivalid;
            amb &= ivalid;
        }
        if (__ballot(amb != 0)) {
            if (amb) {
                const double4 xj = xr[js];
#pragma unroll 1
                for (int k = 0; k < JP_CL; ++k)
                    if ((amb >> k) & 1u)
                        if (exact_in(P, s_xr[w][k], s_im[w][k], xj, ij, sx, sy, sz)) m |= 1u << k;
            }
        }
        if (jc == ci) m &= ~(1u << ja);   // j == i
        uint32_t same = 0;
#pragma unroll
        for (int k = 0; k < JP_CL; ++k) same |= (molj == moli[k] ? 1u : 0u) << k;
        same &= m;
        if (__ballot(same != 0)) {
            for (int k = 0; k < JP_CL; ++k) {
                if (__ballot((same >> k) & 1u) == 0) continue;
                const int a = __builtin_amdgcn_readlane(ai, k);
                const int xb = xptr[a], xe = xptr[a + 1];
                if ((same >> k) & 1u) {
                    for (int e = xb; e < xe; ++e)
                        if (xidx[e] == aj) { m &= ~(1u << k); ++nrem; break; }
                }
            }
        }
#pragma unroll
        for (int k = 0; k < JP_CL; ++k) {
            const int c = __popcll(__ballot((m >> k) & 1u));
            nbv += (l == k) ? c : 0;
        }
        const bool has = m != 0, lj = (tj >> 12) & 1;
        const uint64_t bA = __ballot(has && lj), bB = __ballot(has && !lj);
        const int na = __popcll(bA), nb = __popcll(bB);
        if (nA + nB + na + nb <= cap) {
            const uint2 E = make_uint2((uint32_t)js | (code << JP_SLOT_BITS), m);
            if (has && lj) ent[e0 + nA + lane_rank(bA)] = E;
            if (has && !lj) ent[e0 + cap - 1 - (nB + lane_rank(bB))] = E;
            nA += na;
            nB += nb;
        } else {
            ovf = true;
        }
    };

    // one loop, one drain site: drain while 8 are queued (or the rest at the
    // end), else append the next round of candidate j-clusters
    int qb = -64, t = 0, k0 = 0, k1 = 0;
    uint32_t code = 13;
    double shx = 0.0, shy = 0.0, shz = 0.0;
    bool done = false;
    for (;;) {
        const int nq = qt - qh;
        if (nq >= 8 || (done && nq > 0)) {
            wave_lds_sync();
            drain(min(nq, 8));
            qh += min(nq, 8);
            wave_lds_sync();
            continue;
        }
        if (done) break;
        if (__ballot(k0 + t < k1) == 0) {   // next 64 candidate cells
            qb += 64;
            if (qb >= r.ncand) { done = true; continue; }
            const int q = qb + l;
            k0 = k1 = 0;
            t = 0;
            if (q < r.ncand) {
                const int cell = clu_cell(P, r, crank, cellext, q, code, shx, shy, shz);
                if (cell >= 0) cell_clusters(cslot, ccnt, cell, k0, k1);
            }
            continue;
        }
        const int jc = k0 + t;
        const bool ok = (jc < k1) && clu_near(cbb, jc, shx, shy, shz, r);
        const uint64_t b = __ballot(ok);
        if (ok) s_q[w][(qt + lane_rank(b)) & (JP_QCAP - 1)] = (uint32_t)jc | (code << JP_SLOT_BITS);
        qt += __popcll(b);
        ++t;
    }
    if (l == 0) {
        lcnt[ci] = make_int2(nA, nB);
        if (ovf) atomicOr(err, JP_ERR_LISTCAP);
    }
    if (l < JP_CL && ai >= 0) nbcount[ai] = nbv;
    const int nr = wave_sum(nrem);
    if (l == 0 && nr) atomicAdd(nremoved + s, nr);
}
