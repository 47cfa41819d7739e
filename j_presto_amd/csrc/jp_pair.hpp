// Nonbonded ZD + LJ pair kernels on the 8-atom i-cluster lists
// (enear_ZD_doub_PB semantics: full list, each atom owns its own sum).
//
// One block of PW wavefronts per i-cluster.  The cluster's entries are cut
// into chunks of 64; wave w takes chunks w, w+PW, ...  A chunk is staged
// in the wave's own LDS slice (one entry and its j atom per lane, as
// coordinates relative to the cluster centre with the image shift applied),
// then 8 iterations run over it: lane l computes the pair (i = l & 7,
// j = entry 8*it + (l >> 3)).  Chunks come from the LJ part or the non-LJ part
// of the list, so the LJ term is uniform per chunk.  At the end the 8 lanes
// of each i atom and the PW waves are combined in a fixed order.
//
// FP32 chunks start at multiples of 64 entries: LJ chunks from the front of
// the segment, non-LJ chunks from its end.  The list kernel keeps a byte per
// entry (the 8-bit pair mask) in mb[], with the padding of the two partial
// chunks zeroed, so the masks of one iteration are one 64-bit word of mb
// (bit 8*jsub + ia = pair (ia, entry 8*it + jsub)): a scalar load that selects
// lanes directly, without a per-lane bit test.
#pragma once
#include <type_traits>

#include "jp_list.hpp"

#define JP_PW 4      // default wavefronts per i-cluster in the pair kernels (2, 4 or 8; autotuned)
#define JP_FLUSH_CH 4   // FP32 partial sums go to fixed point every 4 chunks (<= 32 terms)

__device__ __forceinline__ long long fix_f(float v, double scale) { return __double2ll_rn((double)v * scale); }

// Inputs of the ZD excess term, added to the pair sum of each atom in the
// epilogue: gradient = excess (FP64, md_run's order) + pair.
struct ExcessArgs {
    const int *perm;        // slot -> atom (-1: padding)
    const double *x, *chg;
    const int *eptr;
    const int2 *eidx;
    const double *efac;     // per system, per energy class
    __device__ __forceinline__ double grad(const SysP &P, int s, size_t slot, int k) const {
        const int a = perm[slot];
        return a < 0 ? 0.0 : excess_grad_k(a, k, P, x, chg, eptr, eidx, efac + 3 * s);
    }
};

struct ChunkSpan {
    size_t base;
    int n;
    bool lj;
};

// Chunk ch of cluster c in the aligned layout: lanes [lo, hi) of the 64
// entries at base hold entries of the LJ part (front of the segment) or of
// the non-LJ part (back of the segment); base is a multiple of 64 when cap is.
struct ChunkAl {
    size_t base;
    int lo, hi;
    bool lj;
};

__device__ __forceinline__ ChunkAl chunk_al(int ch, size_t e0, int cap, int2 nn) {
    const int chA = (nn.x + 63) >> 6;
    ChunkAl s;
    s.lj = ch < chA;
    if (s.lj) {
        s.base = e0 + 64 * ch;
        s.lo = 0;
        s.hi = min(64, nn.x - 64 * ch);
    } else {
        const int k = ch - chA;
        s.base = e0 + cap - 64 * (k + 1);
        s.lo = max(0, 64 * (k + 1) - nn.y);
        s.hi = 64;
    }
    return s;
}

// v where bit (lane) of the wave mask m is set, else 0 (one v_cndmask_b32 on
// an SGPR-pair mask).
__device__ __forceinline__ float lane_select(float v, uint64_t m) {
    float r;
    asm("v_cndmask_b32 %0, 0, %1, %2" : "=v"(r) : "v"(v), "s"(m));
    return r;
}

// Chunk ch of cluster c: entries [base, base+n) of the LJ part (front of the
// segment) or of the non-LJ part (back of the segment).
__device__ __forceinline__ ChunkSpan chunk_span(int ch, size_t e0, int cap, int2 nn) {
    const int chA = (nn.x + 63) >> 6;
    ChunkSpan s;
    s.lj = ch < chA;
    const int off = s.lj ? 64 * ch : 64 * (ch - chA);
    s.n = min(64, (s.lj ? nn.x : nn.y) - off);
    s.base = s.lj ? e0 + off : e0 + cap - nn.y + off;
    return s;
}

// FP32 path: pair arithmetic in FP32 (coordinates relative to the cluster
// centre, |r| < ~20 A), per-lane partial sums flushed to 64-bit fixed point
// every <= 32 pairs: deterministic and independent of the launch shape.
template <bool ENERGY, bool LMD, bool SMALL, int PW>
__global__ void __launch_bounds__(64 * PW)
k_pair_f32(int nclu, const int *__restrict__ clu_sys, const SysP *__restrict__ sp, const double4 *__restrict__ ccen,
           const JAtom *__restrict__ ys, const float2 *__restrict__ lj32, int ccap,
           const int2 *__restrict__ lcnt, const uint2 *__restrict__ ent, const uint64_t *__restrict__ mb,
           const ExcessArgs ex, double *__restrict__ gslot, long long *__restrict__ epart) {
    extern __shared__ float2 ljf[];
    __shared__ double s_sh[27][3];
    __shared__ float4 s_pos[PW][64];
    __shared__ int s_meta[PW][64];
    __shared__ long long s_red[PW][JP_CL][3];
    __shared__ long long s_ered[PW][6];
    const int c = xcd_remap(blockIdx.x, gridDim.x);
    if (c >= nclu) return;   // whole block
    const SysP &P = sp[clu_sys[c]];
    const int tid = threadIdx.x, l = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6), ia = l & 7, jsub = l >> 3;
    const double4 cen = ccen[c];
    // force steps use (6 A, 12 B) so that the LJ force factor is (12 B/r^6 - 6 A) / r^8
    for (int k = tid; k < P.ntyp * P.ntyp; k += 64 * PW) {
        const float2 ab = lj32[P.lj0 + k];
        ljf[k] = ENERGY ? ab : make_float2(6.f * ab.x, 12.f * ab.y);
    }
    if (tid < 27) {
        int sx, sy, sz;
        shift_of((uint32_t)tid, sx, sy, sz);
        s_sh[tid][0] = P.L[0] * sx - cen.x;
        s_sh[tid][1] = P.L[1] * sy - cen.y;
        s_sh[tid][2] = P.L[2] * sz - cen.z;
    }
    const JAtom ai = ys[c * JP_CL + ia];
    const float xi = (float)(ai.x - cen.x), yi = (float)(ai.y - cen.y), zi = (float)(ai.z - cen.z);
    const float qi = ai.q;
    const int trow = (ai.t & 0xfff) * P.ntyp, cli = (ai.t >> 13) & 1;
    float f0 = 1.f, f1 = 1.f;
    int k0 = P.cls[0], k1 = P.cls[0];
    if (LMD) {
        f0 = (float)P.fac[cli * 2]; f1 = (float)P.fac[cli * 2 + 1];
        k0 = P.cls[cli * 2]; k1 = P.cls[cli * 2 + 1];
    }
    const float rc2 = (float)P.rc2, fcf = (float)P.fcoeff, bcf = (float)P.bcoeff, zcf = (float)P.zcore;
    const float Lx = (float)P.L[0], Ly = (float)P.L[1], Lz = (float)P.L[2];
    const float iLx = (float)P.invL[0], iLy = (float)P.invL[1], iLz = (float)P.invL[2];
    const size_t e0 = (size_t)c * ccap;
    const int cap = ccap;
    const int2 nn = lcnt[c];
    const int nch = ((nn.x + 63) >> 6) + ((nn.y + 63) >> 6);
    __syncthreads();

    float fx = 0.f, fy = 0.f, fz = 0.f;
    float hx = 0.f, hy = 0.f, hz = 0.f;   // chunks without LJ (force steps): sum of q_j c d, times q_i at the flush
    long long ax = 0, ay = 0, az = 0;
    float ee[3] = {0.f, 0.f, 0.f}, ev[3] = {0.f, 0.f, 0.f};
    long long eacc[6] = {0, 0, 0, 0, 0, 0};
    auto flush = [&]() {
        if (!ENERGY) {
            fx += qi * hx; fy += qi * hy; fz += qi * hz;
            hx = hy = hz = 0.f;
        }
        ax += fix_f(fx, JP_FSCALE); ay += fix_f(fy, JP_FSCALE); az += fix_f(fz, JP_FSCALE);
        fx = fy = fz = 0.f;
        if (ENERGY)
            for (int k = 0; k < 3; ++k) {
                eacc[k] += fix_f(ee[k], JP_ESCALE); eacc[3 + k] += fix_f(ev[k], JP_ESCALE);
                ee[k] = ev[k] = 0.f;
            }
    };
    // force steps: pair masks from mb (scalar), no per-lane bit test; s_meta
    // holds the byte offset 8*type_j of j's LJ column (| class << 20 with LMD)
    const char *ljrow = reinterpret_cast<const char *>(ljf + trow);
    auto pair_m = [&](auto ljtag, uint64_t M, int e) {
        constexpr bool LJ = decltype(ljtag)::value;
        const float4 pj = s_pos[w][e];
        float dx = xi - pj.x, dy = yi - pj.y, dz = zi - pj.z;
        if (SMALL) {
            dx = dx - Lx * rintf(dx * iLx);
            dy = dy - Ly * rintf(dy * iLy);
            dz = dz - Lz * rintf(dz * iLz);
        }
        const float r2 = dx * dx + dy * dy + dz * dz;
        // lanes outside M (r2 may be 0 there) are removed by the select,
        // never by arithmetic
        const float ri = __builtin_amdgcn_rsqf(r2);
        const float ri2 = ri * ri;
        const uint64_t in = __builtin_amdgcn_ballot_w64(r2 <= rc2) & M;
        if (!LJ) {
            float cq = pj.w * (ri * ri2 - fcf);
            if (LMD) cq *= ((s_meta[w][e] >> 20) & 1) ? f1 : f0;
            cq = lane_select(cq, in);
            hx += cq * dx;
            hy += cq * dy;
            hz += cq * dz;
        } else {
            const int mt = s_meta[w][e];
            const float2 q6 = *reinterpret_cast<const float2 *>(ljrow + (LMD ? (mt & 0xfffff) : mt));
            const float ri6 = ri2 * ri2 * ri2;
            const float lj = (q6.y * ri6 - q6.x) * ri6;
            float cf = (qi * pj.w) * (ri * ri2 - fcf) + lj * ri2;
            if (LMD) cf *= ((mt >> 20) & 1) ? f1 : f0;
            cf = lane_select(cf, in);
            fx += cf * dx;
            fy += cf * dy;
            fz += cf * dz;
        }
    };
    auto run_m = [&](auto ljtag, const uint64_t *mrow, int it0, int it1) {
        if (it0 == 0 && it1 == 8) {   // full chunk
#pragma unroll
            for (int it = 0; it < 8; ++it) pair_m(ljtag, mrow[it], it * 8 + jsub);
        } else {
            for (int it = it0; it < it1; ++it) pair_m(ljtag, mrow[it], it * 8 + jsub);
        }
    };
    auto run = [&](auto ljtag, int it0, int it1) {
        constexpr bool LJ = decltype(ljtag)::value;
        for (int it = it0; it < it1; ++it) {
            const float4 pj = s_pos[w][it * 8 + jsub];
            const int mt = s_meta[w][it * 8 + jsub];
            float dx = xi - pj.x, dy = yi - pj.y, dz = zi - pj.z;
            if (SMALL) {   // an edge < 2*rlim: re-take the minimum image (enear.f90:189-191)
                dx = dx - Lx * rintf(dx * iLx);
                dy = dy - Ly * rintf(dy * iLy);
                dz = dz - Lz * rintf(dz * iLz);
            }
            const float r2 = dx * dx + dy * dy + dz * dz;
            const bool in = ((mt >> ia) & 1) && (r2 <= rc2);
            // pairs not in the list or beyond rc (r2 may be 0 there) are removed by
            // the selects below, never by arithmetic
            const float ri = __builtin_amdgcn_rsqf(r2);
            const float ri2 = ri * ri;
            const float qq = qi * pj.w;
            float cf = qq * (ri * ri2 - fcf);
            float vv = 0.f;
            if (LJ) {
                const float2 q6 = ljf[trow + ((mt >> 8) & 0xfff)];
                // 1/r^2 correctly rounded (IEEE division) rather than the rsq
                // estimate squared: r^-12 would carry twelve times its error, and the
                // near-cancelling LJ energy sums it over every pair of the system
                const float rr2 = ENERGY ? 1.0f / r2 : ri2;
                const float ri6 = rr2 * rr2 * rr2;
                const float v1 = q6.x * ri6, v2 = q6.y * ri6 * ri6;
                cf += (12.f * v2 - 6.f * v1) * rr2;
                vv = v2 - v1;
            }
            if (LMD) cf *= ((mt >> 20) & 1) ? f1 : f0;
            cf = in ? cf : 0.f;
            fx += cf * dx;
            fy += cf * dy;
            fz += cf * dz;
            if (ENERGY) {
                const float el = in ? qq * (ri + bcf * r2 - zcf) : 0.f;
                vv = in ? vv : 0.f;
                const int k = ((mt >> 20) & 1) ? k1 : k0;
                ee[0] += (k == 0) ? el : 0.f; ee[1] += (k == 1) ? el : 0.f; ee[2] += (k == 2) ? el : 0.f;
                ev[0] += (k == 0) ? vv : 0.f; ev[1] += (k == 1) ? vv : 0.f; ev[2] += (k == 2) ? vv : 0.f;
            }
        }
    };
    // Software pipeline over this wave's chunks: while a chunk computes, the
    // j atoms of the next chunk and the entries of the one after are in flight.
    // The loads are unconditional (lanes without an entry read entry e0 and
    // j slot 0, both allocated, and ignore them), so that no wait is placed
    // where they are issued.
    auto entry_of = [&](int ch, uint2 &E) -> bool {
        bool ok = false;
        size_t idx = e0;
        if (ch < nch) {
            const ChunkAl s = chunk_al(ch, e0, cap, nn);
            ok = l >= s.lo && l < s.hi;
            if (ok) idx = s.base + l;
        }
        E = ent[idx];
        return ok;
    };
    uint2 Ec, En;
    bool okc = entry_of(w, Ec), okn = entry_of(w + PW, En);
    JAtom bc = ys[okc ? (Ec.x & JP_SLOT_MASK) : 0u];
    int nfl = 0;
    for (int ch = w; ch < nch; ch += PW) {
        const ChunkAl sp_ = chunk_al(ch, e0, cap, nn);
        wave_lds_sync();
        if (okc) {
            const int code = (int)(Ec.x >> JP_SLOT_BITS);
            s_pos[w][l] = make_float4((float)(bc.x + s_sh[code][0]), (float)(bc.y + s_sh[code][1]),
                                      (float)(bc.z + s_sh[code][2]), bc.q);
            s_meta[w][l] = ENERGY ? (int)(Ec.y & 0xffu) | ((bc.t & 0xfff) << 8) | (((bc.t >> 13) & 1) << 20)
                                  : ((bc.t & 0xfff) << 3) | (((bc.t >> 13) & 1) << 20);
        } else {
            s_pos[w][l] = make_float4(0.f, 0.f, 0.f, 0.f);
            s_meta[w][l] = 0;
        }
        wave_lds_sync();
        const JAtom bn = ys[okn ? (En.x & JP_SLOT_MASK) : 0u];
        uint2 Enn;
        const bool oknn = entry_of(ch + 2 * PW, Enn);
        const int it0 = sp_.lo >> 3, it1 = (sp_.hi + 7) >> 3;
        if (ENERGY) {
            if (sp_.lj) run(std::true_type{}, it0, it1);
            else run(std::false_type{}, it0, it1);
        } else {
            const uint64_t *mrow = mb + (sp_.base >> 3);
            if (sp_.lj) run_m(std::true_type{}, mrow, it0, it1);
            else run_m(std::false_type{}, mrow, it0, it1);
        }
        if (++nfl == JP_FLUSH_CH) { nfl = 0; flush(); }
        Ec = En; okc = okn; bc = bn;
        En = Enn; okn = oknn;
    }
    flush();
#pragma unroll
    for (int o = 8; o < 64; o <<= 1) {
        ax += __shfl_xor(ax, o, 64);
        ay += __shfl_xor(ay, o, 64);
        az += __shfl_xor(az, o, 64);
    }
    if (jsub == 0) { s_red[w][ia][0] = ax; s_red[w][ia][1] = ay; s_red[w][ia][2] = az; }
    if (ENERGY)
        for (int k = 0; k < 6; ++k) {
            const long long t = wave_sum(eacc[k]);
            if (l == 0) s_ered[w][k] = t;
        }
    __syncthreads();
    if (tid < JP_CL * 3) {
        const int i = tid / 3, k = tid % 3;
        long long t = 0;
        for (int q = 0; q < PW; ++q) t += s_red[q][i][k];
        const size_t sl = (size_t)c * JP_CL + i;
        {
#pragma clang fp contract(off)
            gslot[3 * sl + k] = ex.grad(P, clu_sys[c], sl, k) + (double)(-t) * (1.0 / JP_FSCALE);
        }
    } else if (ENERGY && tid >= 32 && tid < 38) {
        long long t = 0;
        for (int q = 0; q < PW; ++q) t += s_ered[q][tid - 32];
        epart[6 * (size_t)c + tid - 32] = t;
    }
}

// Plain FP64 path: md_run's per-pair arithmetic (enear.f90:182-258 ZD lines)
// in FP64 with IEEE division and square root and no FMA contraction.
template <bool ENERGY, bool LMD, bool SMALL, int PW>
__global__ void __launch_bounds__(64 * PW)
k_pair_f64(int nclu, const int *__restrict__ clu_sys, const SysP *__restrict__ sp, const double4 *__restrict__ ccen,
           const JAtom *__restrict__ ys, const double *__restrict__ qd, const double2 *__restrict__ lj64,
           int ccap, const int2 *__restrict__ lcnt, const uint2 *__restrict__ ent,
           const ExcessArgs ex, double *__restrict__ gslot, double *__restrict__ epart) {
#pragma clang fp contract(off)
    extern __shared__ double2 ljs[];
    __shared__ double s_sh[27][3];
    __shared__ double4 s_pos[PW][64];
    __shared__ int s_meta[PW][64];
    __shared__ double s_red[PW][JP_CL][3];
    __shared__ double s_ered[PW][6];
    const int c = xcd_remap(blockIdx.x, gridDim.x);
    if (c >= nclu) return;
    const SysP &P = sp[clu_sys[c]];
    const int tid = threadIdx.x, l = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6), ia = l & 7, jsub = l >> 3;
    const double4 cen = ccen[c];
    for (int k = tid; k < P.ntyp * P.ntyp; k += 64 * PW) ljs[k] = lj64[P.lj0 + k];
    if (tid < 27) {
        int sx, sy, sz;
        shift_of((uint32_t)tid, sx, sy, sz);
        s_sh[tid][0] = P.L[0] * sx - cen.x;
        s_sh[tid][1] = P.L[1] * sy - cen.y;
        s_sh[tid][2] = P.L[2] * sz - cen.z;
    }
    const int islot = c * JP_CL + ia;
    const JAtom ai = ys[islot];
    const double xi = ai.x - cen.x, yi = ai.y - cen.y, zi = ai.z - cen.z;
    const double qi = qd[islot];
    const int trow = (ai.t & 0xfff) * P.ntyp, cli = (ai.t >> 13) & 1;
    double f0 = 1.0, f1 = 1.0;
    int k0 = P.cls[0], k1 = P.cls[0];
    if (LMD) {
        f0 = P.fac[cli * 2]; f1 = P.fac[cli * 2 + 1];
        k0 = P.cls[cli * 2]; k1 = P.cls[cli * 2 + 1];
    }
    const size_t e0 = (size_t)c * ccap;
    const int cap = ccap;
    const int2 nn = lcnt[c];
    const int nch = ((nn.x + 63) >> 6) + ((nn.y + 63) >> 6);
    __syncthreads();

    double gx = 0.0, gy = 0.0, gz = 0.0;
    double ee[3] = {0.0, 0.0, 0.0}, ev[3] = {0.0, 0.0, 0.0};
    auto run = [&](auto ljtag, int nit) {
#pragma clang fp contract(off)
        constexpr bool LJ = decltype(ljtag)::value;
        for (int it = 0; it < nit; ++it) {
            const double4 pj = s_pos[w][it * 8 + jsub];
            const int mt = s_meta[w][it * 8 + jsub];
            double dx = xi - pj.x, dy = yi - pj.y, dz = zi - pj.z;
            if (SMALL) {
                dx = dx - P.L[0] * anint_d(dx * P.invL[0]);
                dy = dy - P.L[1] * anint_d(dy * P.invL[1]);
                dz = dz - P.L[2] * anint_d(dz * P.invL[2]);
            }
            double dd2 = dx * dx + dy * dy + dz * dz;
            const bool in = (mt >> ia) & 1;
            dd2 = in ? dd2 : 1.0;
            const double dd2i = 1.0 / dd2;
            double co = (dd2i - P.irlim2 >= 0.0) ? 1.0 : 0.0;
            co = in ? co : 0.0;
            const double ddi = sqrt(dd2i);
            const double eel = qi * pj.w * co;
            double cc = eel * (dd2i * ddi - P.fcoeff);
            double vv = 0.0;
            if (LJ) {
                const double2 q6 = ljs[trow + ((mt >> 8) & 0xfff)];
                const double dd6i = dd2i * dd2i * dd2i * co;
                const double v1 = q6.x * dd6i;
                const double v2 = q6.y * dd6i * dd6i;
                vv = v2 - v1;
                cc = cc + (-6.0 * v1 + 12.0 * v2) * dd2i;
            }
            if (LMD) cc = cc * (((mt >> 20) & 1) ? f1 : f0);
            gx = gx + dx * cc;
            gy = gy + dy * cc;
            gz = gz + dz * cc;
            if (ENERGY) {
                const double el = eel * (ddi + P.bcoeff * dd2 - P.zcore);
                const int k = ((mt >> 20) & 1) ? k1 : k0;
                ee[0] += (k == 0) ? el : 0.0; ee[1] += (k == 1) ? el : 0.0; ee[2] += (k == 2) ? el : 0.0;
                ev[0] += (k == 0) ? vv : 0.0; ev[1] += (k == 1) ? vv : 0.0; ev[2] += (k == 2) ? vv : 0.0;
            }
        }
    };
    for (int ch = w; ch < nch; ch += PW) {
        const ChunkSpan sp_ = chunk_span(ch, e0, cap, nn);
        wave_lds_sync();
        if (l < sp_.n) {
            const uint2 E = ent[sp_.base + l];
            const int js = (int)(E.x & JP_SLOT_MASK);
            const int code = (int)(E.x >> JP_SLOT_BITS);
            const JAtom b = ys[js];
            s_pos[w][l] = make_double4(b.x + s_sh[code][0], b.y + s_sh[code][1], b.z + s_sh[code][2], qd[js]);
            s_meta[w][l] = (int)(E.y & 0xffu) | ((b.t & 0xfff) << 8) | (((b.t >> 13) & 1) << 20);
        } else {
            s_pos[w][l] = make_double4(0.0, 0.0, 0.0, 0.0);
            s_meta[w][l] = 0;
        }
        wave_lds_sync();
        const int nit = (sp_.n + 7) >> 3;
        if (sp_.lj) run(std::true_type{}, nit);
        else run(std::false_type{}, nit);
    }
#pragma unroll
    for (int o = 8; o < 64; o <<= 1) {
        gx = gx + __shfl_xor(gx, o, 64);
        gy = gy + __shfl_xor(gy, o, 64);
        gz = gz + __shfl_xor(gz, o, 64);
    }
    if (jsub == 0) { s_red[w][ia][0] = gx; s_red[w][ia][1] = gy; s_red[w][ia][2] = gz; }
    if (ENERGY)
        for (int k = 0; k < 3; ++k) {
            const double a = wave_sum(ee[k]), b = wave_sum(ev[k]);
            if (l == 0) { s_ered[w][k] = a; s_ered[w][3 + k] = b; }
        }
    __syncthreads();
    if (tid < JP_CL * 3) {
        const int i = tid / 3, k = tid % 3;
        double t = 0.0;
        for (int q = 0; q < PW; ++q) t = t + s_red[q][i][k];
        const size_t sl = (size_t)c * JP_CL + i;
        gslot[3 * sl + k] = ex.grad(P, clu_sys[c], sl, k) + (0.0 - t);
    } else if (ENERGY && tid >= 32 && tid < 38) {
        double t = 0.0;
        for (int q = 0; q < PW; ++q) t = t + s_ered[q][tid - 32];
        epart[6 * (size_t)c + tid - 32] = t;
    }
}

// Sums of the per-cluster counts, in cluster order, by one block:
// tot[k] = sum_c part[3c + k].
__global__ void __launch_bounds__(256) k_sum3(int nclu, const unsigned long long *__restrict__ part,
                                              unsigned long long *__restrict__ tot) {
    __shared__ unsigned long long s[3][256];
    unsigned long long a[3] = {0, 0, 0};
    for (int c = threadIdx.x; c < nclu; c += 256)
        for (int k = 0; k < 3; ++k) a[k] += part[3 * (size_t)c + k];
    for (int k = 0; k < 3; ++k) s[k][threadIdx.x] = a[k];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int k = 0; k < 3; ++k) s[k][threadIdx.x] += s[k][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x < 3) tot[threadIdx.x] = s[threadIdx.x][0];
}

// Unique pairs with r <= rc (each pair once) at the current coordinates:
// out[0] all pairs, out[1] pairs where both atoms carry LJ, out[2] listed
// pairs (r <= rlim at the last rebuild, each once).  One wavefront per
// i-cluster.
template <bool PART>
__device__ __forceinline__ void count_pairs_body(int nclu, const int *__restrict__ clu_sys, const SysP *__restrict__ sp,
                                                 const JAtom *__restrict__ ys, int ccap, const int2 *__restrict__ lcnt,
                                                 const uint2 *__restrict__ ent, unsigned long long *__restrict__ out) {
    const int c = blockIdx.x;
    if (c >= nclu) return;
    const SysP &P = sp[clu_sys[c]];
    const int l = threadIdx.x, ia = l & 7, jsub = l >> 3, islot = c * JP_CL + ia;
    const JAtom ai = ys[islot];
    const size_t e0 = (size_t)c * ccap;
    const int cap = ccap;
    const int2 nn = lcnt[c];
    unsigned long long n = 0, nlj = 0, nl = 0;
    for (int e = jsub; e < nn.x + nn.y; e += 8) {
        const uint2 E = ent[e < nn.x ? e0 + e : e0 + cap - nn.y + (e - nn.x)];
        const int js = (int)(E.x & JP_SLOT_MASK);
        int sx, sy, sz;
        shift_of(E.x >> JP_SLOT_BITS, sx, sy, sz);
        const JAtom aj = ys[js];
        double dx = ai.x - (aj.x + P.L[0] * sx), dy = ai.y - (aj.y + P.L[1] * sy), dz = ai.z - (aj.z + P.L[2] * sz);
        if (P.small) {
            dx = dx - P.L[0] * rint(dx * P.invL[0]);
            dy = dy - P.L[1] * rint(dy * P.invL[1]);
            dz = dz - P.L[2] * rint(dz * P.invL[2]);
        }
        const bool lis = ((E.y >> ia) & 1u) && (islot < js);
        const bool in = lis && (dx * dx + dy * dy + dz * dz <= P.rc2);
        n += in;
        nlj += in && ((ai.t >> 12) & 1) && ((aj.t >> 12) & 1);
        nl += lis;
    }
    n = wave_sum(n);
    nlj = wave_sum(nlj);
    nl = wave_sum(nl);
    if (l == 0) {
        if (PART) {
            out[3 * (size_t)c] = n;
            out[3 * (size_t)c + 1] = nlj;
            out[3 * (size_t)c + 2] = nl;
        } else {
            atomicAdd(out, n);
            atomicAdd(out + 1, nlj);
            atomicAdd(out + 2, nl);
        }
    }
}

__global__ void __launch_bounds__(64)
k_count_pairs(int nclu, const int *__restrict__ clu_sys, const SysP *__restrict__ sp, const JAtom *__restrict__ ys,
              int ccap, const int2 *__restrict__ lcnt, const uint2 *__restrict__ ent,
              unsigned long long *__restrict__ out) {
    count_pairs_body<false>(nclu, clu_sys, sp, ys, ccap, lcnt, ent, out);
}

// md_run's list-size check, exactly (cutoff.f90:25-46): per atom and energy
// class, the partners in the vdW table (both atoms carry LJ, itb) and in the
// EL table (the others, itbEL) may not exceed n15mx each.  Runs only when
// some atom has more than n15mx partners in all (the list kernel's cheap
// necessary condition).  One wavefront per i-cluster.
__global__ void __launch_bounds__(64)
k_nbexact(int nclu, const int *__restrict__ clu_sys, const SysP *__restrict__ sp, const JAtom *__restrict__ ys,
          int ccap, const int2 *__restrict__ lcnt, const uint2 *__restrict__ ent, int *__restrict__ flag) {
    __shared__ int s_n[JP_CL][2][3];
    __shared__ int s_t[JP_CL];
    const int c = blockIdx.x;
    if (c >= nclu) return;
    const SysP &P = sp[clu_sys[c]];
    const int l = threadIdx.x;
    if (l < 6 * JP_CL) (&s_n[0][0][0])[l] = 0;
    if (l < JP_CL) s_t[l] = ys[c * JP_CL + l].t;
    __syncthreads();
    const size_t e0 = (size_t)c * ccap;
    const int2 nn = lcnt[c];
    for (int e = l; e < nn.x + nn.y; e += 64) {
        const uint2 E = ent[e < nn.x ? e0 + e : e0 + ccap - nn.y + (e - nn.x)];
        const int tj = ys[E.x & JP_SLOT_MASK].t;
        for (int k = 0; k < JP_CL; ++k) {
            if (!((E.y >> k) & 1u)) continue;
            const int ti = s_t[k];
            const int tab = ((ti >> 12) & 1) & ((tj >> 12) & 1);
            const int cls = P.cls[((ti >> 13) & 1) * 2 + ((tj >> 13) & 1)];
            atomicAdd(&s_n[k][tab][cls], 1);
        }
    }
    __syncthreads();
    if (l < 6 * JP_CL && (&s_n[0][0][0])[l] > P.nblim) atomicOr(flag, 1);
}

// Per-cluster counts of the pairs k_count_pairs counts: out[3c + 0..2] for
// cluster c (plain stores, no atomics).
__global__ void __launch_bounds__(64)
k_count_pairs_part(int nclu, const int *__restrict__ clu_sys, const SysP *__restrict__ sp, const JAtom *__restrict__ ys,
                   int ccap, const int2 *__restrict__ lcnt, const uint2 *__restrict__ ent,
                   unsigned long long *__restrict__ out) {
    count_pairs_body<true>(nclu, clu_sys, sp, ys, ccap, lcnt, ent, out);
}
