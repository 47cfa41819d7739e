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
// Every i-cluster owns a segment of ccap entries at ent[C*ccap]; entries whose
// j carries LJ are stored from its front, the others from its back, so that
// every chunk the pair kernels stage is uniformly with or without the LJ term.
// Each part is kept to whole 64-entry chunks (rounded-up sizes within ccap, a
// multiple of 64), and mb[] holds the mask of every entry as a byte at the
// same index, zero on the padding of the partial chunks.  A segment that would
// overflow is reported (the host grows ccap and builds again).
//
// Candidates: every j-cluster whose home cell (the cell of its first atom) is
// within rlim + 2*ext of C's box, ext bounding how far a cluster's atoms
// stick out of their home cell, then every j-cluster whose own box (shifted by
// s*L) is within rlim of C's box.  Pipeline: k_cbox -> k_list8 (one wavefront
// per i-cluster; the 8 i atoms sit in scalar registers, each lane tests one
// candidate j atom).
//
// The per-pair test runs in FP32 on coordinates about the box centre, taken
// relative to C's centre.  With E the bound on the per-component error (a few
// ulp of 1.5 L), the FP32 d^2 is within 2*rlim*sqrt(3)*E + 4e-7*rlim^2 of the
// FP64 one; pairs with |d^2 - rlim^2| <= band (4x that bound, at least
// 0.01 A^2, SysP::band) are decided by md_run's FP64 test on raw coordinates,
// so the list is the reference's bit for bit.  In boxes with an edge below
// 2*rlim every pair takes the FP64 test, which also checks that s is the
// minimum image.
#pragma once
#include "jp_kernels.hpp"

#define JP_CL 8          // atoms per i-cluster
#define JP_LW 4          // default i-clusters (wavefronts) per block of k_list8 (1, 2, 4 or 8; autotuned)
#define JP_QCAP 128      // per-wave queue of candidate j-clusters (power of 2)
#define JP_LNB 2         // default batches of 8 j-clusters drained per pass of k_list8 (1, 2 or 4; autotuned)

__host__ __device__ __forceinline__ int round64(int n) { return (n + 63) & ~63; }
// list status words lst[]: [0] error bits, [1] longest list above the high-water
// mark, [2 + s] excluded pairs met within rlim by system s
#define JP_ERR_CELL 1
#define JP_ERR_LISTCAP 16     // a segment overflowed: grow ccap and build again
#define JP_ERR_NBLIM 32       // an atom has more listed partners than md_run allows (ier 51)

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

// Box and centres of every cluster (wrapped coordinates at this rebuild: FP64
// box for the candidate search and the pair kernels' frame, FP32 centre for
// the list test), and how far it sticks out of its home cell: maximum per cell
// per cell (non-negative floats compare like their bit patterns; k_crec
// takes the maximum per system).  One lane per slot, the 8 lanes of a
// cluster reduced by shuffles.
__global__ void __launch_bounds__(256)
k_cbox(int nclu, const int *__restrict__ clu_sys, const SysP *__restrict__ sp, const int *__restrict__ perm,
       const int *__restrict__ cellid, const int *__restrict__ ccoord, const JAtom *__restrict__ ys,
       const float4 *__restrict__ lp, double *__restrict__ cbb, double4 *__restrict__ ccen,
       float4 *__restrict__ ccf, unsigned *__restrict__ cellext) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = t >> 3, k = t & 7, l = threadIdx.x & 63;
    const bool act = c < nclu;
    const int s = act ? clu_sys[c] : -1;
    const bool real = act && perm[t] >= 0;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    float flo[3] = {1e30f, 1e30f, 1e30f}, fhi[3] = {-1e30f, -1e30f, -1e30f};
    if (real) {
        const JAtom a = ys[t];
        const float4 q = lp[t];
        lo[0] = hi[0] = a.x; lo[1] = hi[1] = a.y; lo[2] = hi[2] = a.z;
        flo[0] = fhi[0] = q.x; flo[1] = fhi[1] = q.y; flo[2] = fhi[2] = q.z;
    }
#pragma unroll
    for (int o = 1; o < JP_CL; o <<= 1)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            lo[d] = fmin(lo[d], __shfl_xor(lo[d], o, 64));
            hi[d] = fmax(hi[d], __shfl_xor(hi[d], o, 64));
            flo[d] = fminf(flo[d], __shfl_xor(flo[d], o, 64));
            fhi[d] = fmaxf(fhi[d], __shfl_xor(fhi[d], o, 64));
        }
    const int n = __popcll((__ballot(real) >> (l & ~7)) & 0xffull);
    if (act && k == 0) {
        const SysP &P = sp[s];
        for (int d = 0; d < 3; ++d) { cbb[6 * c + d] = lo[d]; cbb[6 * c + 3 + d] = hi[d]; }
        if (n == 0) {
            ccen[c] = make_double4(0.0, 0.0, 0.0, 0.0);
            ccf[c] = make_float4(0.f, 0.f, 0.f, 0.f);
        } else {
            ccen[c] = make_double4(0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1]), 0.5 * (lo[2] + hi[2]), 0.0);
            ccf[c] = make_float4(0.5f * (flo[0] + fhi[0]), 0.5f * (flo[1] + fhi[1]), 0.5f * (flo[2] + fhi[2]), 0.f);
            // slots are real atoms first, padding last: a non-empty cluster's first slot is real
            const int cell = cellid[perm[t]];
            const int cc = ccoord[cell];
            const int cx[3] = {cc & 1023, (cc >> 10) & 1023, (cc >> 20) & 1023};
            double e = 0.0;
            for (int d = 0; d < 3; ++d) {
                const double clo = P.wall[d] + cx[d] * P.csz[d], chi = P.wall[d] + (cx[d] + 1) * P.csz[d];
                e = fmax(e, fmax(clo - lo[d], hi[d] - chi));
            }
            atomicMax(cellext + cell, __float_as_uint((float)e * 1.0001f + 1e-5f));
        }
    }
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
    // candidate indices stay below 2^22 (clu_cell's FP32 division)
    if (!empty && (double)r.n[0] * r.n[1] * r.n[2] >= 4194304.0) r.bad = true;
    r.ncand = (r.bad || empty) ? 0 : r.n[0] * r.n[1] * r.n[2];
    return r;
}

// Per box-grid position: {first slot, atoms, ext bits, 0} of the cell there
// (one load per candidate cell in k_list8), and the largest extent of each
// system (one atomic per wavefront when it holds one system).
__global__ void k_crec(int ncell, const int *__restrict__ crank, const int *__restrict__ cslot,
                       const int *__restrict__ ccnt, const unsigned *__restrict__ cellext, int4 *__restrict__ crec,
                       const int *__restrict__ cell_sys, unsigned *__restrict__ sysext) {
    const int gi = blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = gi < ncell;
    unsigned e = 0u;
    int s = -1;
    if (act) {
        const int c = crank[gi];
        e = cellext[c];
        s = cell_sys[c];
        crec[gi] = make_int4(cslot[c], ccnt[c], (int)e, 0);
    }
    const int s0 = __builtin_amdgcn_readfirstlane(s);
    if (__ballot(act && s != s0) == 0) {
        unsigned m = e;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, o, 64));
        if ((threadIdx.x & 63) == 0 && s0 >= 0) atomicMax(sysext + s0, m);
    } else if (act) {
        atomicMax(sysext + s, e);
    }
}

// Candidate q -> its j-clusters [k0, k1) and image shift; k1 == k0 when the
// unwrapped cell (grown by 2*its ext) is farther than rlim from the cluster
// box.  Shifts of two or more boxes are never within rlim (every edge is
// >= 2*cutoff).
// q / n for 0 <= q < 2^22, 1 <= n < 2^10: FP32 reciprocal estimate, then one
// correction each way (exact).
__device__ __forceinline__ int udiv_small(int q, int n, float rn) {
    int d = (int)((float)q * rn);
    d -= (d * n > q) ? 1 : 0;
    d += ((d + 1) * n <= q) ? 1 : 0;
    return d;
}

__device__ __forceinline__ void clu_cell(const SysP &P, const CluRange &r, const int4 *__restrict__ crec, int q,
                                         uint32_t &code, int &k0, int &k1) {
    k0 = k1 = 0;
    const int q1 = udiv_small(q, r.n[0], 1.0f / (float)r.n[0]);
    const int q2 = udiv_small(q1, r.n[1], 1.0f / (float)r.n[1]);
    const int ux = r.ulo[0] + (q - q1 * r.n[0]), uy = r.ulo[1] + (q1 - q2 * r.n[1]), uz = r.ulo[2] + q2;
    // nc = 2^hbits per edge: floor division is an arithmetic shift
    const int sx = ux >> P.hbits, sy = uy >> P.hbits, sz = uz >> P.hbits;
    if (sx < -1 || sx > 1 || sy < -1 || sy > 1 || sz < -1 || sz > 1) return;
    const int cx = ux - sx * P.nc[0], cy = uy - sy * P.nc[1], cz = uz - sz * P.nc[2];
    const int4 cr = crec[P.cell0 + (cz * P.nc[1] + cy) * P.nc[0] + cx];
    const double m = r.rl + 2.0 * (double)__int_as_float(cr.z);
    const double gx = fmax(0.0, fmax(P.wall[0] + ux * P.csz[0] - r.hi[0], r.lo[0] - (P.wall[0] + (ux + 1) * P.csz[0])));
    const double gy = fmax(0.0, fmax(P.wall[1] + uy * P.csz[1] - r.hi[1], r.lo[1] - (P.wall[1] + (uy + 1) * P.csz[1])));
    const double gz = fmax(0.0, fmax(P.wall[2] + uz * P.csz[2] - r.hi[2], r.lo[2] - (P.wall[2] + (uz + 1) * P.csz[2])));
    if (gx * gx + gy * gy + gz * gz > m * m) return;
    code = (uint32_t)((sx + 1) * 9 + (sy + 1) * 3 + (sz + 1));
    // j-clusters whose first slot lies in this cell
    k0 = (cr.x + JP_CL - 1) / JP_CL;
    k1 = (cr.x + cr.y + JP_CL - 1) / JP_CL;
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

// One wavefront per i-cluster.  Candidate j-clusters go through a queue in
// LDS; every 8 of them are tested as one batch (lane = one j atom against the
// 8 i atoms), NB batches per pass with all of their loads issued before the
// first test (the tile width: more batches keep more loads in flight and use
// more registers).  Per i atom the listed count (nbcount) and per system the number
// of excluded pairs met within rlim (nremoved) are reported for md_run's
// checks.
// HALF: the half list -- only j-clusters after the i-cluster, and inside
// the i-cluster only the pairs (ia < ja): each unique pair once.  Per-atom
// counts then go to icnt (own entries) and jcnt (as a j atom) for k_nbcheck.
template <int LW, bool HALF, int NB>
__global__ void __launch_bounds__(64 * LW) __attribute__((amdgpu_waves_per_eu(NB > 2 ? 4 : 6)))
k_list8(int nclu, const int *__restrict__ clu_sys, const SysP *__restrict__ sp, const double *__restrict__ cbb,
        const float4 *__restrict__ ccf, const int4 *__restrict__ crec, const unsigned *__restrict__ sysext,
        const float4 *__restrict__ lp, const double4 *__restrict__ xr, const int4 *__restrict__ imgs,
        const int *__restrict__ xptr, const int *__restrict__ xidx, const int *__restrict__ mcell,
        const uint64_t *__restrict__ wmask, int ccap, int hiwater, uint2 *__restrict__ ent,
        uint8_t *__restrict__ mb, int2 *__restrict__ lcnt, int *__restrict__ lst, int *__restrict__ icnt,
        int *__restrict__ jcnt) {
    __shared__ uint32_t s_c[LW][JP_QCAP];   // candidate j-clusters of passing cells
    __shared__ uint32_t s_q[LW][JP_QCAP];   // candidates whose box is within rlim
    __shared__ double4 s_xr[LW][JP_CL];
    __shared__ int4 s_im[LW][JP_CL];
    __shared__ float4 s_off[LW][27];        // per image code: L*s - cluster centre (FP32)
    __shared__ int s_mol[LW][JP_CL];
    __shared__ uint64_t s_w[LW][3][JP_CL];  // md_run's cell window of each i atom per axis
    const int l = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // wave-uniform
    const int ci = blockIdx.x * LW + w;
    if (ci >= nclu) return;   // whole wavefront; no block-wide barrier below
    const int s = clu_sys[ci];
    const SysP &P = sp[s];
    // Axes on which md_run's neighbour-cell table drops pairs within rlim.
    // Per such axis the i atoms' md_run cells usually take one or two
    // neighbouring values lo and lo+1: then their windows are two uniform
    // masks and hm (the atoms in cell lo+1) picks between them; a cluster
    // spread wider takes the windows of its 8 atoms from LDS.
    const int ctab = wmask ? P.ctab : 0;
    __shared__ uint64_t s_wlo[LW][3], s_whi[LW][3];
    __shared__ uint32_t s_hm[LW][3];
    bool wide = false;
    if (ctab) {
        const int mc = l < JP_CL ? mcell[ci * JP_CL + l] : 0;
        const bool valid = l < JP_CL && imgs[ci * JP_CL + l].w >= 0;
        if (l < 3 * JP_CL) {
            const int k = l & 7, a = l >> 3;
            const int c = (mcell[ci * JP_CL + k] >> (8 * a)) & 0xff;
            s_w[w][a][k] = c >= 1 ? wmask[(3 * (size_t)s + a) * 64 + c - 1] : ~0ull;
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (!((ctab >> a) & 1)) continue;
            int c = valid ? (mc >> (8 * a)) & 0xff : 255, d = valid ? (mc >> (8 * a)) & 0xff : 0;
#pragma unroll
            for (int o = 1; o < JP_CL; o <<= 1) {
                c = min(c, __shfl_xor(c, o, 64));
                d = max(d, __shfl_xor(d, o, 64));
            }
            const int lo = __builtin_amdgcn_readfirstlane(c), hi = __builtin_amdgcn_readfirstlane(d);
            if (lo > hi) continue;   // no atom
            if (hi - lo > 1) wide = true;
            const uint32_t h = (uint32_t)__ballot(valid && ((mc >> (8 * a)) & 0xff) == hi && hi != lo) & 0xffu;
            if (l == 0) {
                s_hm[w][a] = h;
                s_wlo[w][a] = wmask[(3 * (size_t)s + a) * 64 + lo - 1];
                s_whi[w][a] = wmask[(3 * (size_t)s + a) * 64 + hi - 1];
            }
        }
    }
    const CluRange r = clu_range(P, cbb, ci, __uint_as_float(sysext[s]));
    if (r.bad && l == 0) atomicOr(lst, JP_ERR_CELL);   // runaway coordinates: md_run's ier 28
    const float4 cf = ccf[ci];

    // the 8 i atoms: FP32 coordinates relative to the cluster centre, molecule, atom index
    float fx = 1e15f, fy = 1e15f, fz = 1e15f;
    int mi = -2, ai = -1;
    if (l < JP_CL) {
        const int sl = ci * JP_CL + l;
        const int4 im = imgs[sl];
        ai = im.w;
        if (ai >= 0) {
            const float4 q = lp[sl];
            fx = q.x - cf.x;
            fy = q.y - cf.y;
            fz = q.z - cf.z;
            mi = __float_as_int(q.w) & 0x3fffffff;
        }
        s_xr[w][l] = xr[sl];
        s_im[w][l] = im;
        s_mol[w][l] = mi;
    }
    if (l < 27) {
        int sx, sy, sz;
        shift_of((uint32_t)l, sx, sy, sz);
        s_off[w][l] = make_float4((float)P.L[0] * (float)sx - cf.x, (float)P.L[1] * (float)sy - cf.y,
                                  (float)P.L[2] * (float)sz - cf.z, 0.f);
    }
    // uniform copies (scalar registers); lane k keeps atom k's index and count
    float xi[JP_CL], yi[JP_CL], zi[JP_CL];
#pragma unroll
    for (int k = 0; k < JP_CL; ++k) {
        xi[k] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(fx), k));
        yi[k] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(fy), k));
        zi[k] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(fz), k));
    }
    const uint32_t ivalid = (uint32_t)__ballot(l < JP_CL && ai >= 0);
    // listed partners per i atom: per-lane byte counters (atom k in byte k&3 of
    // cnt[k>>2]), folded into nbv (lane k holds atom k's count) every 255 drains
    int nbv = 0, ndr = 0;
    uint32_t cnt0 = 0, cnt1 = 0;
    auto fold = [&]() {
#pragma unroll
        for (int k = 0; k < JP_CL; ++k) {
            const int c = wave_sum((int)(((k < 4 ? cnt0 : cnt1) >> (8 * (k & 3))) & 255u));
            nbv += (l == k) ? c : 0;
        }
        cnt0 = cnt1 = 0;
        ndr = 0;
    };
    wave_lds_sync();

    const size_t e0 = (size_t)ci * ccap;
    const int cap = ccap;
    const float rlo = (float)(P.rlim2 - P.band), rhi = (float)(P.rlim2 + P.band);
    // any d2 in [rlo, rhi] has |d2 - rlim2| <= band + rounding: a looser trigger
    const float rl2f = (float)P.rlim2, trig = (float)P.band + 0.01f;
    const bool small = P.small != 0;
    int nA = 0, nB = 0, nrem = 0, qh = 0, qt = 0;
    bool ovf = false;
    const int jsub = l >> 3, ja = l & 7;

    // test one queued j-cluster per 8 lanes (lane = atom ja of queue entry qe,
    // valid if qv; q = its lp record, loaded by the caller)
    auto test = [&](bool qv, uint32_t qe, float4 q, int cj) __attribute__((always_inline)) {
        const int jc = (int)(qe & JP_SLOT_MASK);
        const uint32_t code = qe >> JP_SLOT_BITS;
        const int js = jc * JP_CL + ja;
        uint32_t m = 0, amb = 0;
        const int meta = qv ? __float_as_int(q.w) : -1;
        const int molj = meta & 0x3fffffff;
        // j relative to the cluster centre, image applied: 3 roundings like
        // (q - cf) + L*s, within the error bound of the band
        const float4 off = s_off[w][code];
        const float rx = q.x + off.x, ry = q.y + off.y, rz = q.z + off.z;
        float amin = 3.0e38f;
        if (meta >= 0) {
            // bit k = [d2 < rlo]: the sign of d2 - rlo shifted in, atom 7 first
#pragma unroll
            for (int k = JP_CL - 1; k >= 0; --k) {
                const float dx = xi[k] - rx, dy = yi[k] - ry, dz = zi[k] - rz;
                const float d2 = dx * dx + dy * dy + dz * dz;
                m = __builtin_amdgcn_alignbit(m, __float_as_uint(d2 - rlo), 31);
                amin = fminf(amin, fabsf(d2 - rl2f));
            }
            m &= ivalid;
        }
        // pairs near rlim (rare) or in a small box: md_run's FP64 test decides
        if (__ballot(meta >= 0 && (small || amin <= trig))) {
            if (meta >= 0) {
                if (small) {
                    m = 0;
                    amb = 0xffu;
                } else {
#pragma unroll
                    for (int k = 0; k < JP_CL; ++k) {
                        const float dx = xi[k] - rx, dy = yi[k] - ry, dz = zi[k] - rz;
                        const float d2 = dx * dx + dy * dy + dz * dz;
                        amb |= ((d2 >= rlo) && (d2 <= rhi) ? 1u : 0u) << k;
                    }
                }
                amb &= ivalid;
            }
            if (amb) {
                int sx, sy, sz;
                shift_of(code, sx, sy, sz);
                const double4 xj = xr[js];
                const int4 ij = imgs[js];
#pragma unroll 1
                for (int k = 0; k < JP_CL; ++k)
                    if ((amb >> k) & 1u)
                        if (exact_in(P, s_xr[w][k], s_im[w][k], xj, ij, sx, sy, sz)) m |= 1u << k;
            }
        }
        if (jc == ci) m &= HALF ? (1u << ja) - 1u : ~(1u << ja);   // j == i; half: ia < ja only
        // md_run tests only the pairs its neighbour-cell table holds
        // (incel.f90:456-508): without an image shift a pair within rlim is
        // always in it; with one, j's md_run cell must lie in the i atom's
        // window on every shifted axis where the table drops pairs
        if (ctab && __ballot(m != 0 && code != 13u)) {
            if (m != 0 && code != 13u) {
                int sh[3];
                shift_of(code, sh[0], sh[1], sh[2]);
                uint32_t adm = 0xffu;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    if (!((ctab >> a) & 1) || sh[a] == 0) continue;
                    const int c = ((cj >> (8 * a)) & 0xff) - 1;
                    uint32_t b;
                    if (!wide) {
                        const uint32_t h = s_hm[w][a];
                        b = (((s_wlo[w][a] >> c) & 1ull) ? (~h & 0xffu) : 0u) | (((s_whi[w][a] >> c) & 1ull) ? h : 0u);
                    } else {
                        b = 0;
#pragma unroll
                        for (int k = 0; k < JP_CL; ++k) b |= (uint32_t)((s_w[w][a][k] >> c) & 1ull) << k;
                    }
                    adm &= b;
                }
                m &= adm;
            }
        }
        // bit k = [molj == moli[k]]: (molj ^ moli[k]) - 1 is negative only for
        // equal ids (ids < 2^30)
        uint32_t same = 0;
#pragma unroll
        for (int k = JP_CL - 1; k >= 0; --k)
            same = __builtin_amdgcn_alignbit(same, (uint32_t)(molj ^ s_mol[w][k]) - 1u, 31);
        same &= m;
        if (__ballot(same != 0)) {
            const int aj = same ? imgs[js].w : -1;
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
        // bits 0-3 and 4-7 of m spread to bytes: n * 0x204081 puts bit b at 7b + b
        cnt0 += ((m & 15u) * 0x204081u) & 0x01010101u;
        cnt1 += ((m >> 4) * 0x204081u) & 0x01010101u;
        if (++ndr == 255) fold();
        const bool has = m != 0, lj = (meta >> 30) & 1;
        const uint64_t bA = __ballot(has && lj), bB = __ballot(has && !lj);
        const int na = __popcll(bA), nb = __popcll(bB);
        if (round64(nA + na) + round64(nB + nb) <= cap) {
            const uint2 E = make_uint2((uint32_t)js | (code << JP_SLOT_BITS), m);
            const size_t pos = lj ? e0 + nA + lane_rank(bA) : e0 + cap - 1 - (nB + lane_rank(bB));
            if (HALF && has) atomicAdd(jcnt + js, __popc(m));
            if (has) {
                ent[pos] = E;
                mb[pos] = (uint8_t)m;
            }
            nA += na;
            nB += nb;
        } else {
            ovf = true;
        }
    };
    // test j-clusters s_q[qh .. qh+nq) (nq <= 8 * NB) as NB batches of 8, all
    // loaded before the first is tested (unconditional loads: slot 0 for lanes
    // without an entry, so that no wait lands where they are issued)
    auto drain = [&](int nq) {
        bool qv[NB];
        uint32_t qe[NB];
        float4 q[NB];
        int cj[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            qv[b] = 8 * b + jsub < nq;
            qe[b] = qv[b] ? s_q[w][(qh + 8 * b + jsub) & (JP_QCAP - 1)] : 0u;
        }
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int j = qv[b] ? (int)(qe[b] & JP_SLOT_MASK) * JP_CL + ja : 0;
            q[b] = lp[j];
            // md_run cells of the j atoms, only where its cell table drops pairs
            cj[b] = ctab ? mcell[j] : 0;
        }
#pragma unroll
        for (int b = 0; b < NB; ++b)
            if (b == 0 || nq > 8 * b) test(qv[b], qe[b], q[b], cj[b]);
    };

    // One loop, one site per phase, highest priority first: drain 8 * NB
    // queued j-clusters; box-test 64 candidates (independent loads); append
    // one round of the current cell batch's j-clusters (LDS only); test the
    // next 64 candidate cells.  Bounds: the queue holds <= 8 * NB - 1 + 64 and
    // the candidate ring <= 63 + 64 entries when written, both < JP_QCAP.
    static_assert(8 * NB - 1 + 64 < JP_QCAP, "queue bound");
    int ch = 0, ct = 0, qb = -64, rnd = 0, k0 = 0, k1 = 0;
    uint32_t ccode = 13;
    bool cells_done = false;
    for (;;) {
        const int nq = qt - qh, nc = ct - ch;
        if (nq >= 8 * NB || (cells_done && nc == 0 && nq > 0)) {
            wave_lds_sync();
            drain(min(nq, 8 * NB));
            qh += min(nq, 8 * NB);
            wave_lds_sync();
            continue;
        }
        if (nc >= 64 || (cells_done && nc > 0)) {
            const int mc = min(nc, 64);
            bool ok = false;
            uint32_t ce = 0;
            wave_lds_sync();
            if (l < mc) {
                ce = s_c[w][(ch + l) & (JP_QCAP - 1)];
                int sx, sy, sz;
                shift_of(ce >> JP_SLOT_BITS, sx, sy, sz);
                ok = (!HALF || (int)(ce & JP_SLOT_MASK) >= ci) &&
                     clu_near(cbb, (int)(ce & JP_SLOT_MASK), P.L[0] * sx, P.L[1] * sy, P.L[2] * sz, r);
            }
            const uint64_t b = __ballot(ok);
            if (ok) s_q[w][(qt + lane_rank(b)) & (JP_QCAP - 1)] = ce;
            qt += __popcll(b);
            ch += mc;
            continue;
        }
        if (cells_done) break;
        if (__ballot(rnd < k1 - k0) == 0) {   // next 64 candidate cells
            qb += 64;
            if (qb >= r.ncand) { cells_done = true; continue; }
            rnd = 0;
            k0 = k1 = 0;
            if (qb + l < r.ncand) clu_cell(P, r, crec, qb + l, ccode, k0, k1);
            continue;
        }
        const bool p = rnd < k1 - k0;
        const uint64_t b = __ballot(p);
        if (p) s_c[w][(ct + lane_rank(b)) & (JP_QCAP - 1)] = (uint32_t)(k0 + rnd) | (ccode << JP_SLOT_BITS);
        ct += __popcll(b);
        ++rnd;
    }
    // zero masks on the padding of the two partial 64-entry chunks (the
    // FP32 pair kernel reads whole chunks of mb); the parts never share a chunk
    if (nA + l < round64(nA)) mb[e0 + nA + l] = 0;
    if (nB + l < round64(nB)) mb[e0 + cap - 1 - (nB + l)] = 0;
    if (l == 0) {
        lcnt[ci] = make_int2(nA, nB);
        if (ovf) atomicOr(lst, JP_ERR_LISTCAP);
        const int used = round64(nA) + round64(nB);
        if (used > hiwater) atomicMax(lst + 1, used);
    }
    fold();
    if (HALF) {
        if (l < JP_CL) icnt[ci * JP_CL + l] = ai >= 0 ? nbv : 0;
    } else if (l < JP_CL && ai >= 0 && nbv > P.nblim) {
        atomicOr(lst, JP_ERR_NBLIM);
    }
    const int nr = wave_sum(nrem);
    if (l == 0 && nr) atomicAdd(lst + 2 + s, nr);
}
