// FP32 / fixed-point pair kernel on the half list (each unique pair once,
// the i-cluster ordered before the j-cluster; enear_ZD_doub_PB's forces and
// energies, each pair's force applied to both atoms by Newton's third law).
//
// One block of PW wavefronts per i-cluster; wave w takes chunks w, w+PW, ...
// of 64 list entries.  Lane l holds entry l of the chunk -- one j atom with
// its image shift and 8-bit mask -- and runs over the 8 i atoms, whose
// coordinates, charges and types are wave-uniform (scalar registers).  So
// the force on j sums in the lane's own registers (no cross-lane traffic per
// pair) and leaves once per chunk as a 64-bit fixed-point atomic add; the
// force on i atom k sums per lane over the wave's chunks and is reduced once
// per cluster through LDS in a fixed order, then added the same way.
// Integer additions commute, so the gradient does not depend on the order in
// which blocks finish, and a system gives the same bits alone or in a batch.
// k_facc_out then turns the accumulators into the FP64 gradient per slot
// (plus the ZD excess term) and clears them for the next evaluation.
//
// Work skipped per chunk: an i atom none of the 64 entries pairs with (its
// mask bits all zero) costs nothing, and the LJ term runs only for an i atom
// that carries LJ in a chunk whose j atoms carry it.
#pragma once
#include "jp_pair.hpp"

#define JP_PWH 2     // default wavefronts per i-cluster of k_pair_h32 (1, 2 or 4; autotuned)

// floor(v * 2^40) as a 64-bit integer for |v| < 2^23: t = v * 2^8 splits into
// its integer part h and fraction f (t - h is exact in FP32), f * 2^32 is
// exact and truncates to the low word.  Five VALU instructions instead of a
// trip through FP64.
__device__ __forceinline__ unsigned long long fix40(float v) {
    const float t = v * 256.0f;
    const float h = floorf(t);
    const unsigned lo = (unsigned)((t - h) * 4294967296.0f);
    return ((unsigned long long)(unsigned)(int)h << 32) | lo;
}

// floor(v * 2^36) likewise (energies, JP_ESCALE), |v| < 2^27.
__device__ __forceinline__ long long fix36(float v) {
    const float t = v * 16.0f;
    const float h = floorf(t);
    const unsigned lo = (unsigned)((t - h) * 4294967296.0f);
    return (long long)(((unsigned long long)(unsigned)(int)h << 32) | lo);
}

template <bool ENERGY, bool LMD, bool SMALL, int PW>
__global__ void __launch_bounds__(64 * PW)
k_pair_h32(int nclu, const int *__restrict__ clu_sys, const SysP *__restrict__ sp, const double4 *__restrict__ ccen,
           const JAtom *__restrict__ ys, const float2 *__restrict__ lj32, int ccap, const int2 *__restrict__ lcnt,
           const uint2 *__restrict__ ent, unsigned long long *__restrict__ facc, long long *__restrict__ epart) {
    extern __shared__ float2 ljf[];
    __shared__ double s_sh[27][3];
    __shared__ float s_t[PW][3 * JP_CL][64];   // per-lane i-atom force partials, transposed
    __shared__ float s_red[PW][3 * JP_CL];
    __shared__ long long s_ered[PW][6];
    const int c = xcd_remap(blockIdx.x, gridDim.x);
    if (c >= nclu) return;   // whole block
    const SysP &P = sp[clu_sys[c]];
    const int tid = threadIdx.x, l = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const double4 cen = ccen[c];
    // force steps use (6 A, 12 B): the LJ force factor is (12 B/r^6 - 6 A) / r^8
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
    // the 8 i atoms, relative to the cluster centre, in scalar registers
    float xk[JP_CL], yk[JP_CL], zk[JP_CL], qk[JP_CL];
    int tk[JP_CL];
    {
        const JAtom a = ys[c * JP_CL + (l & 7)];
        const float ax = (float)(a.x - cen.x), ay = (float)(a.y - cen.y), az = (float)(a.z - cen.z);
#pragma unroll
        for (int k = 0; k < JP_CL; ++k) {
            xk[k] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ax), k));
            yk[k] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ay), k));
            zk[k] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(az), k));
            qk[k] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(a.q), k));
            tk[k] = __builtin_amdgcn_readlane(a.t, k);
        }
    }
    // class factors and energy classes of (class of i, class of j), LMD
    const float fc0 = (float)P.fac[0], fc1 = (float)P.fac[1], fc2 = (float)P.fac[2], fc3 = (float)P.fac[3];
    const float rc2 = (float)P.rc2, fcf = (float)P.fcoeff, bcf = (float)P.bcoeff, zcf = (float)P.zcore;
    const float Lx = (float)P.L[0], Ly = (float)P.L[1], Lz = (float)P.L[2];
    const float iLx = (float)P.invL[0], iLy = (float)P.invL[1], iLz = (float)P.invL[2];
    const size_t e0 = (size_t)c * ccap;
    const int cap = ccap;
    const int2 nn = lcnt[c];
    const int nch = ((nn.x + 63) >> 6) + ((nn.y + 63) >> 6);
    __syncthreads();

    float fi[JP_CL][3];
#pragma unroll
    for (int k = 0; k < JP_CL; ++k) fi[k][0] = fi[k][1] = fi[k][2] = 0.f;
    float ee[3] = {0.f, 0.f, 0.f}, ev[3] = {0.f, 0.f, 0.f};
    long long eacc[6] = {0, 0, 0, 0, 0, 0};

    // one chunk: lane l holds entry base + l when lo <= l < hi
    auto chunk = [&](auto ljtag, bool ok, uint2 E, const JAtom &b) {
        constexpr bool LJ = decltype(ljtag)::value;
        const int code = ok ? (int)(E.x >> JP_SLOT_BITS) : 13;
        const uint32_t m = ok ? (E.y & 0xffu) : 0u;
        const float xj = (float)(b.x + s_sh[code][0]), yj = (float)(b.y + s_sh[code][1]),
                    zj = (float)(b.z + s_sh[code][2]);
        const float qj = b.q;
        const int clj = (b.t >> 13) & 1;
        const char *ljcol = reinterpret_cast<const char *>(ljf) + 8 * (b.t & 0xfff);
        float fjx = 0.f, fjy = 0.f, fjz = 0.f;
#pragma unroll
        for (int k = 0; k < JP_CL; ++k) {
            const uint64_t mk = __ballot((m >> k) & 1u);
            if (mk == 0) continue;   // wave-uniform: i atom k pairs with no entry of the chunk
            float dx = xk[k] - xj, dy = yk[k] - yj, dz = zk[k] - zj;
            if (SMALL) {   // an edge < 2*rlim: re-take the minimum image (enear.f90:189-191)
                dx = dx - Lx * rintf(dx * iLx);
                dy = dy - Ly * rintf(dy * iLy);
                dz = dz - Lz * rintf(dz * iLz);
            }
            const float r2 = dx * dx + dy * dy + dz * dz;
            // lanes outside `in` (r2 may be 0 there) are removed by the
            // selects, never by arithmetic
            const uint64_t in = __ballot(r2 <= rc2) & mk;
            const float ri = __builtin_amdgcn_rsqf(r2);
            const float ri2 = ri * ri;
            const float qq = qk[k] * qj;
            float cf = qq * (ri * ri2 - fcf);
            float vv = 0.f;
            const int ti = tk[k];
            const bool lji = LJ && ((ti >> 12) & 1);   // wave-uniform
            if (lji) {
                const float2 q6 = *reinterpret_cast<const float2 *>(ljcol + 8 * (ti & 0xfff) * P.ntyp);
                // energy steps: 1/r^2 correctly rounded (the r^-12 term would carry
                // twelve times the rsq estimate's error into the summed energy)
                const float rr2 = ENERGY ? 1.0f / r2 : ri2;
                const float ri6 = rr2 * rr2 * rr2;
                if (ENERGY) {
                    const float v1 = q6.x * ri6, v2 = q6.y * ri6 * ri6;
                    cf += (12.f * v2 - 6.f * v1) * rr2;
                    vv = v2 - v1;
                } else {
                    cf += (q6.y * ri6 - q6.x) * ri6 * rr2;
                }
            }
            int kc = P.cls[0];
            if (LMD) {
                const bool cli = (ti >> 13) & 1;   // wave-uniform
                cf *= clj ? (cli ? fc3 : fc1) : (cli ? fc2 : fc0);
                kc = clj ? (cli ? P.cls[3] : P.cls[1]) : (cli ? P.cls[2] : P.cls[0]);
            }
            cf = lane_select(cf, in);
            fi[k][0] += cf * dx;
            fi[k][1] += cf * dy;
            fi[k][2] += cf * dz;
            fjx -= cf * dx;
            fjy -= cf * dy;
            fjz -= cf * dz;
            if (ENERGY) {
                const float el = lane_select(qq * (ri + bcf * r2 - zcf), in);
                vv = lane_select(vv, in);
                ee[0] += (kc == 0) ? el : 0.f; ee[1] += (kc == 1) ? el : 0.f; ee[2] += (kc == 2) ? el : 0.f;
                ev[0] += (kc == 0) ? vv : 0.f; ev[1] += (kc == 1) ? vv : 0.f; ev[2] += (kc == 2) ? vv : 0.f;
            }
        }
        // the force on j from this cluster (<= 8 pair terms), in fixed point
        if (m != 0u) {
            const size_t js = E.x & JP_SLOT_MASK;
            atomicAdd(facc + 4 * js, fix40(fjx));
            atomicAdd(facc + 4 * js + 1, fix40(fjy));
            atomicAdd(facc + 4 * js + 2, fix40(fjz));
        }
        if (ENERGY)
            for (int q = 0; q < 3; ++q) {
                eacc[q] += fix36(ee[q]);
                eacc[3 + q] += fix36(ev[q]);
                ee[q] = ev[q] = 0.f;
            }
    };
    // Software pipeline over this wave's chunks (as k_pair_f32): the j atoms
    // of the next chunk and the entries of the one after are in flight while a
    // chunk computes; lanes without an entry load entry e0 and slot 0.
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
    for (int ch = w; ch < nch; ch += PW) {
        const bool lj = ch < ((nn.x + 63) >> 6);
        const JAtom bn = ys[okn ? (En.x & JP_SLOT_MASK) : 0u];
        uint2 Enn;
        const bool oknn = entry_of(ch + 2 * PW, Enn);
        if (lj) chunk(std::true_type{}, okc, Ec, bc);
        else chunk(std::false_type{}, okc, Ec, bc);
        Ec = En; okc = okn; bc = bn;
        En = Enn; okn = oknn;
    }
    // i-atom forces: lane partials through LDS, each summed over the 64 lanes
    // in lane order, then over the waves in wave order; one atomic per
    // component and atom
#pragma unroll
    for (int k = 0; k < JP_CL; ++k)
#pragma unroll
        for (int q = 0; q < 3; ++q) s_t[w][3 * k + q][l] = fi[k][q];
    if (ENERGY)
        for (int q = 0; q < 6; ++q) {
            const long long t = wave_sum(eacc[q]);
            if (l == 0) s_ered[w][q] = t;
        }
    __syncthreads();
    if (l < 3 * JP_CL) {
        float sum = 0.f;
        const float4 *row = reinterpret_cast<const float4 *>(&s_t[w][l][0]);
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const float4 v = row[q];
            sum += v.x; sum += v.y; sum += v.z; sum += v.w;
        }
        s_red[w][l] = sum;
    }
    __syncthreads();
    if (tid < 3 * JP_CL) {
        float t = 0.f;
        for (int q = 0; q < PW; ++q) t += s_red[q][tid];
        if (t != 0.f) atomicAdd(facc + 4 * ((size_t)c * JP_CL + tid / 3) + tid % 3, fix40(t));
    } else if (ENERGY && tid >= 32 && tid < 38) {
        long long t = 0;
        for (int q = 0; q < PW; ++q) t += s_ered[q][tid - 32];
        epart[6 * (size_t)c + tid - 32] = t;
    }
}

// Gradient per slot from the fixed-point accumulators of k_pair_h32, plus
// the ZD excess term (FP64, md_run's order); the accumulators are cleared for
// the next evaluation.  Runs after the pair kernel, before anything moves x.
__global__ void __launch_bounds__(256)
k_facc_out(int nslot, const int *__restrict__ slot_sys, const SysP *__restrict__ sp, const ExcessArgs ex,
           unsigned long long *__restrict__ facc, double *__restrict__ gslot) {
#pragma clang fp contract(off)
    const int sl = blockIdx.x * blockDim.x + threadIdx.x;
    if (sl >= nslot) return;
    const int s = slot_sys[sl];
    const SysP &P = sp[s];
    const ulonglong4 a = reinterpret_cast<const ulonglong4 *>(facc)[sl];
    reinterpret_cast<ulonglong4 *>(facc)[sl] = make_ulonglong4(0, 0, 0, 0);
    const long long t[3] = {(long long)a.x, (long long)a.y, (long long)a.z};
    for (int k = 0; k < 3; ++k) gslot[3 * (size_t)sl + k] = ex.grad(P, s, sl, k) + (double)(-t[k]) * (1.0 / JP_FSCALE);
}

// Clears the accumulators (after launches whose results are not used, such
// as timing runs).
__global__ void __launch_bounds__(256) k_facc_zero(int nslot, unsigned long long *__restrict__ facc) {
    const int sl = blockIdx.x * blockDim.x + threadIdx.x;
    if (sl < nslot) reinterpret_cast<ulonglong4 *>(facc)[sl] = make_ulonglong4(0, 0, 0, 0);
}
