// libjpgpu: C ABI (include/jpgpu.h) over the gfx950 kernels in jp_kernels.hpp.
//
// One Engine drives one or many systems ("batch"); a jp_ctx is an Engine of
// one system, a jp_batch an Engine of many.  All systems of a batch share every
// kernel launch: per-atom, per-slot and per-group arrays are concatenated and
// each work item finds its system's parameters (SysP) through an index array.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/jpgpu.h"
#include "jp_pair.hpp"
#include "jp_pair_half.hpp"
#include "jp_bonded.hpp"
#include "jp_cano.hpp"
#include "jp_ge.hpp"
#include "jp_alsd.hpp"
#include "jp_cmm.hpp"
#include "jp_stop.hpp"

#define CK(x)                                                          \
    do {                                                               \
        hipError_t e_ = (x);                                           \
        if (e_ != hipSuccess) {                                        \
            fprintf(stderr, "jpgpu: %s failed: %s (%s:%d)\n", #x,      \
                    hipGetErrorString(e_), __FILE__, __LINE__);        \
            return JP_EHIP_BASE - (int)e_;                             \
        }                                                              \
    } while (0)

namespace {

template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    int alloc(size_t m) {
        if (m <= n && p) return 0;
        if (p) hipFree(p);
        p = nullptr;
        n = 0;
        if (m == 0) m = 1;
        if (hipMalloc(&p, m * sizeof(T)) != hipSuccess) return JP_ENOMEM;
        n = m;
        return 0;
    }
    void release() {
        if (p) hipFree(p);
        p = nullptr;
        n = 0;
    }
};

#define AL(buf, m)                          \
    do {                                    \
        int r_ = (buf).alloc(m);            \
        if (r_) return r_;                  \
    } while (0)

inline unsigned grid1(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

// finish_list(false): the deferred rebuild failed or overflowed; rebuild from
// the coordinates it was meant for (internal, never returned to callers)
#define JP_LIST_REDO (-50)

// Scratch of the entry point that allocates nothing (jp_count_pairs_noalloc):
// three per-cluster pair counts for up to JP_NOALLOC_MAX_CLU i-clusters (2M
// atoms) and the three totals, 6 MB held by every GPU the library runs on
// until the process ends.  Host code reaches it only through the address
// hipGetSymbolAddress returns on each device (noalloc_scratch).
#define JP_NOALLOC_MAX_CLU (1 << 18)
static __device__ __align__(256) unsigned long long scratch[3 * (size_t)JP_NOALLOC_MAX_CLU + 4];
static std::mutex noalloc_mu;
static unsigned long long *noalloc_ptr[64];   // per device

// The device address of scratch on the current device (looked up once per
// device), or nullptr.
static unsigned long long *noalloc_scratch() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    std::lock_guard<std::mutex> g(noalloc_mu);
    if (!noalloc_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        noalloc_ptr[dev] = static_cast<unsigned long long *>(p);
    }
    return noalloc_ptr[dev];
}

// md_run's periodic neighbour-cell table (incel.f90:334-358 and 456-508).
// The list of cell (ix,iy,iz) holds every atom of the cells whose index on
// each axis lies in that axis's window of the cell's own index: i -/+ idelay,
// widened by m?delay where it reaches the box-cell range [m1, m2], wrapped
// into the range modulo iix = m2 - m1 + 1 with a single-precision reciprocal
// (floor(real(k - m1) * (1.0/iix))).  So the table is a product of per-axis
// windows, stored as one 64-bit mask per axis and cell index.  The range
// holds iix cells of rlim/2 although the box edge can be up to two cells
// shorter, so cells facing each other across the periodic boundary can be
// more than the window apart and their pairs are never tested: md_run lists
// no pair between them, even inside the cutoff (box edges of 40.75-47 A,
// 94.5-99.75 A, ...).  P.ctab marks the axes where that drops a pair within
// rlim (per-axis gap between the cells' slabs <= rlim), P.ctab_sym whether the
// drops are symmetric.  A window that leaves [m1, m2] after wrapping, or a
// range outside [1, nv] (the reference writes out of bounds there), leaves
// the table undefined: then no pair is dropped.
static void md_cell_windows(SysP &P, int nv, double siz, double rlim, std::vector<uint64_t> &wm) {
    const size_t base = wm.size();
    wm.resize(base + 3 * 64, 0);
    P.ctab = 0;
    P.ctab_sym = 1;
    if (nv > 64) return;
    const int m = (int)std::ceil(rlim / siz);              // idelay (incel.f90:57)
    const int ic = (int)std::lround(nv * 0.5);             // anint(nv(1)*0.5d0)
    int ctab = 0, sym = 1;
    for (int a = 0; a < 3; ++a) {
        const double h = 0.5 * P.L[a] / siz;
        const int m1 = (int)std::ceil(ic - h), m2 = (int)std::ceil(ic + h);
        const int mdel = (std::ceil(h) - h > 0.5) ? 1 : 2;
        if (m1 < 1 || m2 > nv) return;
        const int ii = m2 - m1 + 1;
        const float ri = 1.0f / (float)ii;
        uint64_t *W = wm.data() + base + 64 * a;
        for (int c = m1; c <= m2; ++c) {
            int j1 = c - m, j2 = c + m;
            if (j1 <= m1) j1 -= mdel;
            if (j2 >= m2) j2 += mdel;
            for (int k = j1; k <= j2; ++k) {
                const int kk = k - ii * (int)std::floor((float)(k - m1) * ri);
                if (kk < m1 || kk > m2) return;
                W[c - 1] |= 1ull << (kk - 1);
            }
        }
        // slab of cell c on this axis, clipped to the box
        auto slab = [&](int c, double &lo, double &hi) {
            lo = std::max(P.rorg[a] + (c - 1) * siz, P.wall[a]);
            hi = std::min(P.rorg[a] + c * siz, P.wall[a] + P.L[a]);
        };
        for (int c = m1; c <= m2; ++c)
            for (int d = m1; d <= m2; ++d) {
                const bool cd = (W[c - 1] >> (d - 1)) & 1, dc = (W[d - 1] >> (c - 1)) & 1;
                if (cd && dc) continue;
                double alo, ahi, blo, bhi;
                slab(c, alo, ahi);
                slab(d, blo, bhi);
                if (!(alo < ahi) || !(blo < bhi)) continue;   // a cell no atom can be in
                double gap = 1e300;
                for (int t = -1; t <= 1; ++t) {
                    const double sh = t * P.L[a];
                    gap = std::min(gap, std::max(0.0, std::max(blo + sh - ahi, alo - (bhi + sh))));
                }
                if (gap <= rlim * (1.0 + 1e-12)) {
                    ctab |= 1 << a;
                    if (cd != dc) sym = 0;
                }
            }
    }
    P.ctab = ctab;
    P.ctab_sym = sym;
}

struct HostSys {
    int n = 0, ntyp = 0, nfrag = 1, nfrg = 1, lmd = 0;
    double lambda = 0.0, lambda_v = 0.0, dself[3] = {0, 0, 0};
    double dself_zd[3] = {0, 0, 0};   // the ZD self energy (dself is 0 under CMM)
    double cutoff = 0.0, ccoef = 0.0;  // CUTLEN; 332.06378 / DIEVAL
    int nexcl = 0;
    int n15mx = 0;
};

struct Engine {
    int nsys = 0, natom = 0, nslot = 0, nclu = 0, ncell = 0;
    int n2 = 0, n3 = 0, n4 = 0;  // SHAKE groups of 2, 3 and 4 atoms
    int prec_built = -1;
    bool have_list = false;
    double shake_tol = 1e-6;
    int shake_maxit = 1000;
    std::vector<HostSys> hs;
    std::vector<SysP> sp;
    std::vector<int> atom_sys_h, slot_sys_h, cell_sys_h, clu_sys_h;
    std::vector<int> atom_cls_h;    // cluster (icls - 1) of every atom
    std::vector<double> mass_h;
    std::vector<double> charge_h;   // raw charges (fxchrg)
    hipStream_t st = nullptr;
    hipEvent_t ev[8];               // [0,1] run, [2,3] list build, [4,5] host-launched pair kernels
    bool pe_pending = false;        // ev[4,5] bracket a pair launch not yet added to t_pair
    int n_captures = 0;             // step graphs captured so far
    int captures_last_run = 0;      // of them, during the last run_nve
    double t_total = 0, t_pair = 0, t_list = 0, t_step = 0;
    int n_pair = 0, n_list = 0, n_pair_steps = 0;   // timed pair launches, steps of the last run
    long long idloop = 0;          // md.f90 idloop: steps done, kept across run calls
    int ccap = 0;                   // entries per i-cluster segment of the current list
    int ccap_want = 2048;           // segment size for the next rebuild (from the last one's maximum)
    int ccap_force = 0;             // JP_OPT_LIST_SEGMENT: exact segment size of the next rebuild

    // device arrays
    DBuf<SysP> d_sp;
    DBuf<int> d_atom_sys, d_slot_sys, d_cell_sys, d_crank, d_ccoord, d_clu_sys;
    DBuf<double> d_x, d_v, d_xold, d_vback, d_mass, d_imass, d_chg, d_grad, d_eexc, d_ekin;
    DBuf<int> d_tflag, d_mol, d_xptr, d_xidx, d_eptr;
    DBuf<int2> d_eidx;
    DBuf<double> d_efac;
    DBuf<double2> d_lj64;
    DBuf<float2> d_lj32;
    DBuf<int> d_subkey;
    DBuf<int> d_img, d_cellid, d_refcell, d_cnt, d_start, d_fill, d_inum, d_perm, d_cslot;
    DBuf<long long> d_tmp;    // per cell, (sub-cell << 32 | atom) keys in arrival order
    DBuf<JAtom> d_ys;
    DBuf<double> d_qd, d_cbb;
    DBuf<double4> d_xr, d_ccen;
    DBuf<int> d_mcell;              // md_run cell of every slot, 8 bits per axis
    DBuf<uint64_t> d_wmask;         // md_run's neighbour-cell windows: [system][axis][cell - 1], bit c - 1
    std::vector<uint64_t> wmask_h;
    int cell_table = 1;             // JP_OPT_CELL_TABLE: lists follow md_run's neighbour-cell table
    DBuf<unsigned long long> d_facc;  // k_pair_h32's fixed-point gradient accumulators, [slot][4]
    DBuf<int> d_icnt, d_jcnt;       // half lists: listed partners per slot as i and as j
    DBuf<int> d_mcnt;               // atoms per md_run cell and cluster class, all systems
    DBuf<int> d_nbflag;             // k_nbexact's result
    int nmcnt = 0;
    int half_opt = 0;               // JP_OPT_HALF_LIST: FP32 runs on the half list where it can hold md_run's list
    bool list_half = false;         // the current list is a half list
    // The FP32 path takes the half list unless md_run's cell table drops pairs
    // within rlim asymmetrically in some system (a half list holds a pair
    // either both ways or not at all).
    bool want_half(int prec) const {
        if (cmm || !half_opt || prec != JP_PREC_FP32_FIXED) return false;
        for (int s = 0; s < nsys; ++s)
            if (cell_table && sp[s].ctab && !sp[s].ctab_sym) return false;
        return true;
    }
    // the list the pair kernel of `prec` runs on
    int ensure_list(int prec) {
        if (cmm) return have_list ? 0 : build_list();
        if (have_list && list_half == want_half(prec)) return 0;
        list_half = want_half(prec);
        return build_list();
    }
    DBuf<float4> d_lp, d_ccf;
    DBuf<int4> d_imgs;
    DBuf<int> d_err, d_lst;   // SHAKE error bits; list status words (jp_list.hpp)
    DBuf<unsigned> d_cellext, d_sysext;
    DBuf<int2> d_lcnt;
    DBuf<int4> d_crec;
    DBuf<uint2> d_ent;
    DBuf<uint64_t> d_mb;            // pair mask of every entry, one byte each (8 per word)
    DBuf<double> d_gslot, d_ep64;
    DBuf<long long> d_epfix;
    DBuf<int> d_s2sys, d_s3sys;
    DBuf<int4> d_s2a, d_s3a, d_s4a;
    DBuf<double> d_s2d, d_s3d, d_s4d;
    DBuf<int> d_s4sys;
    DBuf<unsigned long long> d_u64;
    // bonded terms (jp_bonded.hpp): bonds, angles, torsions (+ impropers),
    // 1-4 pairs, per-atom gather lists
    int nbt = 0, nat = 0, ntt = 0, n14 = 0, nbat = 0;
    int nunit = 0;             // k_step units: SHAKE groups and free atoms
    DBuf<int4> d_unit;
    std::vector<int4> units_h;
    // METHOD= CANO (jp_set_canonical): temperature groups and the state of the
    // SHAKE/temperature rounds
    bool cano = false;
    int cano_icslop = 3, ngrp = 0, nab = 0, nub = 0;
    double cano_fcstol = 1e-3;
    DBuf<int> d_gid, d_gsys, d_sg0, d_sg1, d_ab0, d_ab1, d_ub0, d_ub1, d_sfirst, d_shkon, d_cdone;
    DBuf<int2> d_ablk, d_ublk;
    DBuf<double> d_gdof, d_ctemps, d_fcc, d_xcb, d_pa, d_pu, d_pn;
    int set_canonical(int on, double temps, int icslop, double fcstol, const int *group);
    // md_run's heating loop (jp_set_heating: HEATLO, STARTT): while md_run's
    // step count idloop is at most heat_lp the canonical target temperature
    // is heat_t0 + ((temps - heat_t0) / heat_lp) idloop (verlethe.f90:419-423)
    // STOPCE (jp_set_stopce): per system {mode, first atom, atoms stopped,
    // atoms of atom 1's molecule kind}; mode 1 TRAN, 2 ROTA, 3 BOTH
    int stop_mode = 0;
    std::vector<int4> stop_h;
    DBuf<int4> d_stop;
    int set_stopce(int mode, const int *natom, const int *nkind);
    void stop_dof(std::vector<double> &dof, const std::vector<long long> &keys, const int *group) const;
    int start_step(int prec);
    int heat_lp = 0;
    double heat_t0 = 0.0, cano_temps = 300.0, ttgt_set = -1.0;
    DBuf<double> d_ttgt;                // the target of the step being issued, per system
    int set_heating(int htlp, double tmpsta);
    double heat_target(long long id) const {
        if (heat_lp <= 0 || id > heat_lp) return cano_temps;
        return heat_t0 + ((cano_temps - heat_t0) * (1.0 / (double)heat_lp)) * (double)id;
    }
    bool heating(long long id) const { return cano && heat_lp > 0 && id <= heat_lp; }
    int put_target(double t);
    CanoArgs cano_args();
    int cano_step();
    bool cano_retry() const { return cano && nsys == 1 && (n2 + n3 + n4) > 0; }
    int save_state(double *x, double *v, int *e);
    int load_state(const double *x, const double *v, const int *e);
    int shake_failed(bool &failed);
    int cano_retry_step(int prec);
    int cano_steps_eager(int prec, int n);
    // METHOD= GE (jp_set_mcmd): multicanonical weight of every step's forces
    bool ge = false;
    int ge_nw = 0, ge_nd = 0;
    DBuf<double> d_gepar, d_gew;   // parameters; weight and class-1 energy of the step
    DBuf<int> d_eslot;             // energy slot of every bonded energy term
    int set_mcmd(int on, int nw, int nd, const double *c, const double *low, const double *high, double alphalw,
                 double alphaup, double celw, double ceup, double tempce);
    int ge_weight(int prec);
    // ALSD (jp_set_alsd): lambda as a dynamical variable (jp_alsd.hpp)
    bool alsd = false;
    int al_nw = 0, al_nd = 0, al_tctl = 1;
    DBuf<double> d_alpar, d_lam, d_lamb, d_lamb2, d_lmass, d_dself;   // parameters, state, two backups
    DBuf<int> d_lgrp, d_esys, d_ecls;
    std::vector<int> lgrp_h;
    std::vector<double> lmass_h;
    bool cano_grouped = false;          // jp_set_canonical got temperature groups (caltmp2)
    std::vector<double> gdof_h;         // degrees of freedom per group, without lambda's
    std::vector<int> sg0_h, sg1_h;
    int set_alsd(int on, int nw, int nd, const double *c, const double *low, const double *high, double lower,
                 double upper, double celw, double ceup, double tempce, double lmass, int tctl);
    int alsd_force(int prec);
    int alsd_fac();
    DBuf<double> d_lambk;               // lambda's state beside d_xbak (a deferred rebuild's redo)
    // CAL15M= CMM (jp_set_cmm, jp_cmm.hpp): the cell multipole method in
    // place of ZD for one non-periodic system, FP64 whatever the precision
    // asked for.  Slots are atoms (inum = identity, no images).
    bool cmm = false;
    int cmm_upd = 5;                    // CMMUPD: far field every cmm_upd steps (md.f90:679-684)
    bool cmm_itwin = true;              // md_run's itwin: the next step computes the far field
    bool cmm_far_now = true;            // the far field of the force evaluation being issued
    CmmArgs cm{};
    DBuf<double> d_cmm_cg, d_cmm_c0, d_cmm_q, d_cmm_sx0, d_cmm_pot, d_cmm_gnb, d_cmm_enb, d_cmm_gwk, d_cmm_efar,
        d_cmm_eatom, d_cmm_wb, d_cmm_wb2, d_cmm_gwkp;
    DBuf<double4> d_cmm_sx;
    DBuf<int> d_cmm_key, d_cmm_kcnt, d_cmm_kstart, d_cmm_kfill, d_cmm_tmp, d_cmm_sorig, d_cmm_spos, d_cmm_sflag,
        d_cmm_smol, d_cmm_npc, d_cmm_err, d_cmm_xwide, d_cmm_l2cnt, d_cmm_l2ptr, d_cmm_l2idx, d_cmm_vflag,
        d_cmm_vpos, d_cmm_vlist, d_cmm_kvcnt, d_cmm_kvstart, d_cmm_vf, d_cmm_va, d_cmm_vm, d_cmm_vp, d_cmm_tsum;
    DBuf<double4> d_cmm_vx0;
    DBuf<unsigned long long> d_cmm_xmask;
    int cmm_scan(int n, const int *in, int *out);
    int cmm_vcap = 0;                   // k_cmm_lj2's grid: atoms with a second-nearest table, only grows
    int set_cmm(int on, double cminsiz, int nlev, int cmmupd);
    int cmm_celprp();
    int cmm_forces(bool energy, bool itwin);
    int cmm_sums();
    int cmm_time(int reps, double &near, double &far);
    // the far field of the step after `k` steps of a segment that starts now
    bool cmm_due(int k) const { return k == 0 ? cmm_itwin : (idloop + k) % cmm_upd == 0; }
    // non-bonded forces of the current method into d_gslot
    int nb(int prec, bool energy, bool itwin);
    int lam_upload();                   // host lambda and velocity -> device state
    int lam_download();                 // device state -> host lambda and velocity
    // energy sums on the device: per-system atom and cluster ranges, results,
    // and a pinned host copy of them
    DBuf<int> d_abeg, d_aend, d_cbeg, d_cend;
    DBuf<double> d_rsum;        // [nsys][6] pair (FP64), [nsys][3] excess, [nsys] kinetic
    DBuf<long long> d_rfix;     // [nsys][6] pair (fixed point)
    DBuf<double> d_part;        // k_seg_part partials
    DBuf<long long> d_partf;
    // per-system sums of a[i * ncomp + k] over segments [beg, end) into out
    template <typename T>
    int seg_sum(const int *beg, const int *end, int ncomp, const T *a, T *part, T *out, int nseg = -1) {
        if (nseg < 0) nseg = nsys;
        hipLaunchKernelGGL(k_seg_part<T>, dim3(JP_SEG_NB, nseg), 256, 0, st, beg, end, ncomp, a, part);
        hipLaunchKernelGGL(k_seg_fin<T>, nseg, 64, 0, st, ncomp, (const T *)part, out);
        CK(hipGetLastError());
        return 0;
    }
    double *h_rsum = nullptr;
    long long *h_rfix = nullptr;
    int *h_err = nullptr;
    // per energy entry of k_bonded_energy: system, class, energy slot (0..5 =
    // bond, angle, torsion, improper, 1-4 vdW, 1-4 elec)
    std::vector<int> esys_h, ecls_h, eslot_h;
    DBuf<int> d_bat, d_bptr, d_bref, d_torc;
    DBuf<int4> d_bnd, d_ang, d_tor, d_p14;
    DBuf<double2> d_bndp, d_angp, d_p14p;
    DBuf<double4> d_torp;
    DBuf<double> d_eterm;

    ~Engine() {
        drop_graphs();
        if (st) {
            hipStreamSynchronize(st);
            for (auto &e : ev) hipEventDestroy(e);
            hipStreamDestroy(st);
        }
        d_sp.release(); d_atom_sys.release(); d_slot_sys.release();
        d_cell_sys.release(); d_crank.release(); d_ccoord.release(); d_clu_sys.release(); d_x.release(); d_v.release(); d_xold.release(); d_vback.release();
        d_mass.release(); d_imass.release(); d_chg.release(); d_grad.release(); d_eexc.release(); d_ekin.release();
        d_tflag.release(); d_mol.release(); d_xptr.release(); d_xidx.release(); d_eptr.release(); d_eidx.release();
        d_efac.release(); d_lj64.release(); d_lj32.release(); d_img.release(); d_cellid.release();
        d_refcell.release(); d_subkey.release(); d_cnt.release(); d_start.release(); d_fill.release(); d_tmp.release(); d_inum.release();
        d_perm.release(); d_cslot.release(); d_ys.release(); d_qd.release(); d_cbb.release(); d_xr.release();
        d_ccen.release(); d_imgs.release(); d_lp.release(); d_ccf.release(); d_err.release(); d_lst.release();
        d_cellext.release(); d_sysext.release(); d_lcnt.release(); d_crec.release();
        d_ent.release(); d_mb.release(); d_gslot.release(); d_ep64.release(); d_epfix.release();
        d_s2sys.release(); d_s3sys.release(); d_s2a.release(); d_s3a.release(); d_s2d.release(); d_s3d.release();
        d_s4sys.release(); d_s4a.release(); d_s4d.release();
        d_u64.release();
        d_bat.release(); d_bptr.release(); d_bref.release(); d_bnd.release(); d_ang.release();
        d_bndp.release(); d_angp.release(); d_eterm.release(); d_unit.release();
        d_gid.release(); d_gsys.release(); d_sg0.release(); d_sg1.release(); d_ab0.release(); d_ab1.release();
        d_ub0.release(); d_ub1.release(); d_ablk.release(); d_ublk.release(); d_sfirst.release(); d_shkon.release(); d_cdone.release(); d_stop.release();
        d_gdof.release(); d_ctemps.release(); d_ttgt.release(); d_fcc.release(); d_xcb.release(); d_pa.release(); d_pu.release(); d_pn.release();
        d_torc.release(); d_tor.release(); d_p14.release(); d_p14p.release(); d_torp.release();
        d_abeg.release(); d_aend.release(); d_cbeg.release(); d_cend.release(); d_rsum.release(); d_rfix.release();
        d_part.release(); d_partf.release();
        d_gepar.release(); d_gew.release(); d_eslot.release();
        d_alpar.release(); d_lam.release(); d_lamb.release(); d_lamb2.release(); d_lmass.release(); d_dself.release();
        d_lgrp.release(); d_esys.release(); d_ecls.release(); d_lambk.release();
        d_cmm_cg.release(); d_cmm_c0.release(); d_cmm_q.release(); d_cmm_sx0.release(); d_cmm_pot.release();
        d_cmm_gnb.release(); d_cmm_enb.release(); d_cmm_gwk.release(); d_cmm_efar.release(); d_cmm_eatom.release();
        d_cmm_wb.release(); d_cmm_wb2.release(); d_cmm_gwkp.release(); d_cmm_sx.release(); d_cmm_key.release(); d_cmm_kcnt.release();
        d_cmm_kstart.release(); d_cmm_kfill.release(); d_cmm_tmp.release(); d_cmm_sorig.release();
        d_cmm_spos.release(); d_cmm_sflag.release(); d_cmm_smol.release(); d_cmm_npc.release(); d_cmm_err.release();
        d_cmm_xwide.release(); d_cmm_l2cnt.release(); d_cmm_l2ptr.release(); d_cmm_l2idx.release();
        d_cmm_vflag.release(); d_cmm_vpos.release(); d_cmm_vlist.release(); d_cmm_xmask.release();
        d_cmm_kvcnt.release(); d_cmm_kvstart.release(); d_cmm_vf.release(); d_cmm_va.release(); d_cmm_vm.release();
        d_cmm_vp.release(); d_cmm_tsum.release(); d_cmm_vx0.release();
        if (h_rsum) hipHostFree(h_rsum);
        if (h_rfix) hipHostFree(h_rfix);
        if (h_err) hipHostFree(h_err);
        if (h_lst) hipHostFree(h_lst);
        d_xbak.release(); d_vbak.release(); d_errbak.release();
    }

    int init(const jp_sys_desc *descs, int ns);
    int upload_sp() {
        CK(hipMemcpyAsync(d_sp.p, sp.data(), sizeof(SysP) * nsys, hipMemcpyHostToDevice, st));
        return 0;
    }
    // celprp.  defer = true: enqueue the rebuild and its status read-back and
    // return; finish_list() later reads the status (retry = false: an
    // overflow or error is returned as JP_LIST_REDO for the caller to rebuild
    // from restored coordinates).
    int build_list(bool defer = false);
    int finish_list(bool retry);
    bool nblim_exceeded();
    int cell_err = 0;          // k_bin's flag survives a retry
    int *h_lst = nullptr;      // pinned list status words
    DBuf<double> d_xbak, d_vbak;   // state before a step segment that overlaps a list check
    // METHOD= CANO, one system: state before a segment and before a step of
    // verlethe's SHAKE retry (its dt halving after a SHAKE failure)
    DBuf<double> d_xs, d_vs, d_xs2, d_vs2;
    DBuf<int> d_errs;
    int n_retry = 0;               // steps that needed the retry in the last run
    DBuf<int> d_errbak;
    int refresh_sorted();
    ExcessArgs excess_args() const {
        ExcessArgs a;
        a.perm = d_perm.p; a.x = d_x.p; a.chg = d_chg.p; a.eptr = d_eptr.p; a.eidx = d_eidx.p; a.efac = d_efac.p;
        return a;
    }
    int pair(int prec, bool energy);
    int pair_out(int prec);          // half list: accumulators -> gradient slots (after pair())
    int facc_zero();
    int pair_variant() const;
    BondedArgs bonded_args() const {
        BondedArgs b;
        b.nbat = nbat; b.bat = d_bat.p; b.bptr = d_bptr.p; b.bref = d_bref.p;
        b.bnd = d_bnd.p; b.ang = d_ang.p; b.bndp = d_bndp.p; b.angp = d_angp.p;
        b.tor = d_tor.p; b.torp = d_torp.p; b.torc = d_torc.p; b.p14 = d_p14.p; b.p14p = d_p14p.p;
        b.tflag = d_tflag.p; b.lj64 = d_lj64.p;
        return b;
    }
    int forces(int prec, bool energy);   // pair kernel, then the bonded terms into the same slots
    int bonded_energies(std::vector<double> &et);   // [nsys][3 classes][6 slots]
    int bonded(const double *cord, double *grad, double *eterm);
    int set_factors(int onehot_class);  // -1: run factors (lambda), k: one-hot class k
    int nonbond(const double *cord, double *grad, double *eelec, double *evdw, int prec);
    int energies(int prec, std::vector<double> &eel, std::vector<double> &evd, std::vector<double> &eex);
    int shake(int &ier_out);
    int run_nve(int nsteps, int update, int nmon, int prec, jp_energy_log *log);
    int restart_step(int prec, jp_energy_log *log);
    int prepare_graphs(int nsteps, int update, int nmon, int prec);
    int integrate(int prec, bool mon);   // kick + SHAKE + v update of one step (after the forces)
    int mon_record(int step, int nsteps, int nmon, int nlog, const std::vector<double> &eel,
                   const std::vector<double> &evd, const std::vector<double> &et, jp_energy_log *log);

    // Steps between list rebuilds and energy steps run as captured HIP graphs:
    // nsteps x (pair, kick, SHAKE, post), cached per shape.  Pair-kernel
    // times come from host-launched steps only (t_pair / n_pair).
    struct StepGraph {
        hipGraphExec_t exec = nullptr;
        int nsteps = 0, prec = -1, ccap = -1, variant = -1;
        int cmm_key = -1;   // CMM: far-field phase of the segment's first step, 2 * (idloop % CMMUPD) + itwin
        const void *ent = nullptr;
    };
    std::vector<StepGraph> graphs;
    int use_graphs = 1;
    // launch shapes: wavefronts per i-cluster (pair kernels) and per block
    // (k_list8); autotune = 1 times the candidates at the next jp_run_nve
    int pw = JP_PW, lw = JP_LW, lnb = JP_LNB, autotune = 0, tuned = 0;
    int grid_shift = 0;             // JP_OPT_GRID_SHIFT: box-grid cells 2^shift times finer than 4.5-9 A
    int grid_tables();
    int regrid(int shift);
    float last_build_ms = 0.f;      // device time of the last list rebuild
    int pwh = JP_PWH;               // wavefronts per i-cluster of the half-list force kernel
    int tune(int prec, int update);
    StepGraph *step_graph(int n, int prec);
    void drop_graphs();
    int get_list(int32_t *cell3, int64_t *npairs, int32_t *ptr, int32_t *idx, int64_t cap);
    int count_pairs(int64_t *np, int64_t *nlj, int64_t *nlist);
    int count_pairs_noalloc(int64_t *np, int64_t *nlj, int64_t *nlist);
    int time_pair(int prec, int reps, double &sec);
};

static void launch_list8(Engine &g, int hiwater);

// The ier of the device status word of a step: SHAKE's, then remotv's.
static int step_ier(int err) {
    if (err & 8) return JP_IER_SHAKE_SING;
    if (err & 4) return JP_IER_SHAKE_CONV;
    if (err & JP_ERR_STOP_MASS) return JP_IER_STOP_MASS;
    if (err & JP_ERR_STOP_SING) return JP_IER_STOP_SING;
    return 0;
}

// Box grid of one system: 2^hbits cells per edge, box-aligned, binned in
// Hilbert order.  shift 0 gives cells of 4.5-9 A; each unit of shift halves
// (positive) or doubles (negative) the cell edge, within 1 <= hbits <= 9 (the
// 2x2x2 sub-cell keys of k_bin hold 3 * (hbits + 1) bits).
static void set_grid(SysP &P, int shift) {
    const double lmin = std::min(P.L[0], std::min(P.L[1], P.L[2]));
    const int hb0 = std::min(9, std::max(1, (int)std::floor(std::log2(lmin / 4.5))));
    const int hb = std::min(9, std::max(1, hb0 + shift));
    P.hbits = hb;
    for (int k = 0; k < 3; ++k) {
        P.nc[k] = 1 << hb;
        P.csz[k] = P.L[k] / P.nc[k];
        P.icsz[k] = P.nc[k] / P.L[k];
    }
}

int Engine::init(const jp_sys_desc *descs, int ns) {
    if (ns <= 0) return JP_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return JP_ENODEV;
    CK(hipStreamCreate(&st));
    for (auto &e : ev) CK(hipEventCreate(&e));
    if (!noalloc_scratch()) return JP_ENODEV;   // set-up of jp_count_pairs_noalloc's scratch on this device
    nsys = ns;
    hs.resize(ns);
    sp.resize(ns);
    // host-side concatenation
    std::vector<double> x, chg, mass, imass;
    std::vector<int> tflag, mol, xptr{0}, xidx, eptr{0};
    std::vector<int2> eidx;
    std::vector<double2> lj64;
    std::vector<float2> lj32;
    std::vector<int> s2sys, s3sys, s4sys;
    std::vector<int4> s2a, s3a, s4a;
    std::vector<double> s2d, s3d, s4d;
    std::vector<int4> bnd, ang, tor, p14;
    std::vector<double2> bndp, angp, p14p;
    std::vector<double4> torp;
    std::vector<int> torc;
    std::vector<std::vector<int>> bref_of;   // per global atom: term codes in kind order
    // energy-entry metadata per kind (system, class, slot), concatenated after the loop
    std::vector<int> em_sys[4], em_cls[4], em_slot[4];
    int atom0 = 0, slot0 = 0, lj0 = 0;
    shake_tol = descs[0].shake_tol > 0 ? descs[0].shake_tol : 1e-6;
    shake_maxit = descs[0].shake_maxit > 0 ? descs[0].shake_maxit : 1000;
    for (int s = 0; s < ns; ++s) {
        const jp_sys_desc &d = descs[s];
        HostSys &h = hs[s];
        SysP &P = sp[s];
        memset(&P, 0, sizeof(P));
        const int n = d.natom, nt = d.ntyp;
        if (n <= 0 || nt <= 0 || nt > 64 || d.nfrag < 1 || d.nfrag > 2 || d.nfrg < 1 || d.nfrg > 3) return JP_EINVAL;
        h.n = n; h.ntyp = nt; h.nfrag = d.nfrag; h.nfrg = d.nfrg; h.lmd = d.lmd; h.lambda = d.lambda;
        h.nexcl = d.nexcl;
        h.cutoff = d.cutoff;
        h.ccoef = 332.06378 / d.diel;
        // box (inpmd.f90:578-583, 764-768) and cell grids
        double rc = d.cutoff, rlim = rc + 1.5;
        for (int k = 0; k < 3; ++k) {
            P.L[k] = d.celwal[2 * k + 1] - d.celwal[2 * k];
            P.invL[k] = 1.0 / P.L[k];
            P.wall[k] = d.celwal[2 * k];
            if (!(P.L[k] >= 2.0 * rc)) return JP_EINVAL;  // inpmd.f90:694-698
        }
        set_grid(P, grid_shift);
        // md_run's cell numbering (incel.f90:29-57, mk_c, celpar.f90:93-108)
        {
            int nlev = d.cell_level > 0 ? d.cell_level : 3;
            double cminsiz = rlim * 0.5, ext = 0.0;
            for (int k = 0; k < 3; ++k) {
                double lo = d.x[k], hi = d.x[k];
                for (int i = 1; i < n; ++i) { lo = std::min(lo, d.x[3 * i + k]); hi = std::max(hi, d.x[3 * i + k]); }
                ext = std::max(ext, hi - lo);
            }
            if ((double)(1 << (nlev + 1)) * cminsiz < ext) nlev = (int)std::ceil(std::log(ext / cminsiz) / std::log(2.0) - 1.0);
            int nv = 4 << (nlev - 1);
            double siz = cminsiz, r = -0.5 * (nv + 1), zz = siz * 0.5;
            for (int k = 0; k < 3; ++k) {
                double fxcbou = (d.celwal[2 * k] + d.celwal[2 * k + 1]) * 0.5;
                double c11 = siz * (1.0 + r) + fxcbou;
                P.rorg[k] = c11 - zz;
            }
            P.rr = 1.0 / siz;
            md_cell_windows(P, nv, siz, rlim, wmask_h);
            P.mnv = nv;
            P.ipmax = (int)std::ceil(siz * siz * siz * 0.27) * 2;   // incel.f90:146 (PARALV= DOUB)
            P.mc0 = nmcnt;
            nmcnt += 2 * nv * nv * nv;
        }
        P.rlim2 = rlim * rlim;
        {
            // FP32 list-test band (jp_list.hpp): E = 3 ulp of 1.5 * the longest edge
            const double lmax = std::max(P.L[0], std::max(P.L[1], P.L[2]));
            const double ulp = std::ldexp(1.0, (int)std::floor(std::log2(1.5 * lmax)) - 23);
            const double E = 3.0 * ulp;
            P.band = std::max(0.01, 4.0 * (2.0 * rlim * std::sqrt(3.0) * E + 4e-7 * rlim * rlim));
            for (int k = 0; k < 3; ++k) P.ctr[k] = P.wall[k] + 0.5 * P.L[k];
        }
        P.rc2 = rc * rc;
        P.small = (P.L[0] < 2.0 * rlim || P.L[1] < 2.0 * rlim || P.L[2] < 2.0 * rlim) ? 1 : 0;
        P.irlim2 = 1.0 / (rc * rc);
        double rt = 1.0 / rc;
        P.zcore = 1.5 * rt;
        P.fcoeff = rt * rt * rt;
        P.bcoeff = 0.5 * P.fcoeff;
        P.dt = d.dt;
        P.atom0 = atom0; P.natom = n;
        int npad = (n + 63) / 64 * 64;
        P.slot0 = slot0; P.nslot = npad;
        P.lj0 = lj0; P.ntyp = nt;
        h.n15mx = (int)std::ceil(0.15 * (4.0 / 3.0 * 3.14159265358979 * rc * rc * rc)) * 2;
        P.nblim = h.n15mx;
        // energy class of each cluster pair (n_matrix is column-major, 1-based)
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) {
                int aa = std::min(a, d.nfrag - 1), bb = std::min(b, d.nfrag - 1);
                P.cls[a * 2 + b] = d.n_matrix[bb * d.nfrag + aa] - 1;
                P.fac[a * 2 + b] = 1.0;
            }
        // LJ table, AMBER combination (inptpl.f90:397-419)
        std::vector<double> c6((size_t)nt * nt), c12((size_t)nt * nt);
        for (int i = 0; i < nt; ++i)
            for (int j = i; j < nt; ++j) {
                double rtmp = std::sqrt(d.eps[i] * d.eps[j]);
                double r2 = d.rmin2[i] + d.rmin2[j];
                r2 = r2 * r2;
                r2 = r2 * r2 * r2;
                double a6 = rtmp * r2 * 2.0, a12 = rtmp * r2 * r2;
                c6[i * nt + j] = c6[j * nt + i] = a6;
                c12[i * nt + j] = c12[j * nt + i] = a12;
            }
        for (int k = 0; k < nt * nt; ++k) {
            lj64.push_back(make_double2(c6[k], c12[k]));
            lj32.push_back(make_float2((float)c6[k], (float)c12[k]));
        }
        std::vector<char> zt(nt);
        for (int t = 0; t < nt; ++t) {   // input.f90:58-69
            bool z = true;
            for (int j = 0; j < nt; ++j)
                if (std::fabs(c6[j * nt + t]) >= 1e-16 || std::fabs(c12[j * nt + t]) >= 1e-16) z = false;
            zt[t] = z;
        }
        // atoms
        double ccoef = 332.06378 / d.diel, fac = std::sqrt(ccoef);   // inpcls.f90:215-223
        for (int i = 0; i < n; ++i) {
            for (int k = 0; k < 3; ++k) x.push_back(d.x[3 * i + k]);
            chg.push_back(fac * d.charge[i]);
            charge_h.push_back(d.charge[i]);
            mass.push_back(d.mass[i]);
            imass.push_back(1.0 / d.mass[i]);
            int t = d.atype[i] - 1, c = d.icls[i] - 1;
            if (t < 0 || t >= nt || c < 0 || c >= d.nfrag) return JP_EINVAL;
            tflag.push_back(t | ((zt[t] ? 0 : 1) << 12) | (c << 13));
            atom_sys_h.push_back(s);
            atom_cls_h.push_back(c);
        }
        // molecules = connected components of the excluded-pair graph
        std::vector<int> par(n);
        std::iota(par.begin(), par.end(), 0);
        auto find = [&](int a) { while (par[a] != a) { par[a] = par[par[a]]; a = par[a]; } return a; };
        std::vector<std::vector<int>> ex(n);
        for (int e = 0; e < d.nexcl; ++e) {
            int a = d.excl[2 * e] - 1, b = d.excl[2 * e + 1] - 1;
            if (a < 0 || b < 0 || a >= n || b >= n) return JP_EINVAL;
            ex[a].push_back(b + atom0);
            ex[b].push_back(a + atom0);
            par[find(a)] = find(b);
        }
        for (int i = 0; i < n; ++i) {
            mol.push_back(find(i) + atom0);
            std::sort(ex[i].begin(), ex[i].end());
            xidx.insert(xidx.end(), ex[i].begin(), ex[i].end());
            xptr.push_back((int)xidx.size());
        }
        // excess pairs (incel.f90:629-688): per-atom partner lists in pair order
        std::vector<std::vector<int2>> exs(n);
        std::vector<double> dwork((size_t)n * 3, 0.0);
        auto clsof = [&](int a, int b) { return d.n_matrix[(d.icls[b] - 1) * d.nfrag + (d.icls[a] - 1)] - 1; };
        for (int i = 0; i < n; ++i) dwork[(size_t)i * 3 + clsof(i, i)] += d.charge[i];
        for (int e = 0; e < d.nexcess; ++e) {
            int a = d.excess[2 * e] - 1, b = d.excess[2 * e + 1] - 1;
            if (a < 0 || b < 0 || a >= n || b >= n) return JP_EINVAL;
            int k = clsof(a, b);
            exs[a].push_back(make_int2(b + atom0, k));
            exs[b].push_back(make_int2(a + atom0, k));
            dwork[(size_t)a * 3 + k] += d.charge[b];
            dwork[(size_t)b * 3 + k] += d.charge[a];
        }
        for (int i = 0; i < n; ++i) {
            eidx.insert(eidx.end(), exs[i].begin(), exs[i].end());
            eptr.push_back((int)eidx.size());
        }
        for (int k = 0; k < d.nfrg; ++k) {
            double sum = 0.0;
            for (int i = 0; i < n; ++i) sum = sum + d.charge[i] * dwork[(size_t)i * 3 + k];
            h.dself[k] = -P.zcore * sum * 0.5 * ccoef;
            h.dself_zd[k] = h.dself[k];
        }
        // SHAKE groups
        for (int g = 0; g < d.nshake; ++g) {
            int na = d.shake_natom[g];
            const int32_t *aa = d.shake_atoms + 4 * g;
            int4 q = make_int4(aa[0] - 1 + atom0, aa[1] - 1 + atom0, na > 2 ? aa[2] - 1 + atom0 : -1, -1);
            for (int k = 0; k < na; ++k)
                if (aa[k] < 1 || aa[k] > n) return JP_EINVAL;
            if (na == 2) {
                s2a.push_back(q); s2sys.push_back(s);
                for (int k = 0; k < 6; ++k) s2d.push_back(d.shake_dist2[6 * g + k]);
            } else if (na == 3) {
                s3a.push_back(q); s3sys.push_back(s);
                for (int k = 0; k < 6; ++k) s3d.push_back(d.shake_dist2[6 * g + k]);
            } else if (na == 4) {   // shake4_PB (shake.f90:333-562)
                q.w = aa[3] - 1 + atom0;
                s4a.push_back(q); s4sys.push_back(s);
                for (int k = 0; k < 6; ++k) s4d.push_back(d.shake_dist2[6 * g + k]);
            } else {
                return JP_EINVAL;
            }
        }
        // bonded terms (dnergv.f90:62-193); class default max(n_matrix) (incel.f90:952-969)
        {
            int cmax = 0;
            for (int q = 0; q < d.nfrag * d.nfrag; ++q) cmax = std::max(cmax, d.n_matrix[q] - 1);
            if (d.nbond < 0 || d.nangle < 0) return JP_EINVAL;
            if ((d.nbond && (!d.bond_atoms || !d.bond_fc || !d.bond_r0)) ||
                (d.nangle && (!d.angle_atoms || !d.angle_fc || !d.angle_theta0)))
                return JP_EINVAL;
            bref_of.resize((size_t)atom0 + n);
            for (int b = 0; b < d.nbond; ++b) {
                const int i = d.bond_atoms[2 * b] - 1, j = d.bond_atoms[2 * b + 1] - 1;
                const int c = d.bond_cls ? d.bond_cls[b] - 1 : cmax;
                if (i < 0 || j < 0 || i >= n || j >= n || i == j || c < 0 || c >= d.nfrg) return JP_EINVAL;
                const int t = (int)bnd.size();
                bnd.push_back(make_int4(i + atom0, j + atom0, c, 0));
                bndp.push_back(make_double2(d.bond_fc[b], d.bond_r0[b]));
                em_sys[0].push_back(s); em_cls[0].push_back(c); em_slot[0].push_back(0);
                bref_of[i + atom0].push_back(bref_code(t, JP_BK_BOND, 0));
                bref_of[j + atom0].push_back(bref_code(t, JP_BK_BOND, 1));
            }
            for (int g = 0; g < d.nangle; ++g) {
                const int i = d.angle_atoms[3 * g] - 1, j = d.angle_atoms[3 * g + 1] - 1, k = d.angle_atoms[3 * g + 2] - 1;
                const int c = d.angle_cls ? d.angle_cls[g] - 1 : cmax;
                if (i < 0 || j < 0 || k < 0 || i >= n || j >= n || k >= n || i == j || j == k || i == k || c < 0 ||
                    c >= d.nfrg)
                    return JP_EINVAL;
                const int t = (int)ang.size();
                ang.push_back(make_int4(i + atom0, j + atom0, k + atom0, c));
                angp.push_back(make_double2(d.angle_fc[g], d.angle_theta0[g]));
                em_sys[1].push_back(s); em_cls[1].push_back(c); em_slot[1].push_back(1);
                bref_of[i + atom0].push_back(bref_code(t, JP_BK_ANGLE, 0));
                bref_of[j + atom0].push_back(bref_code(t, JP_BK_ANGLE, 1));
                bref_of[k + atom0].push_back(bref_code(t, JP_BK_ANGLE, 2));
            }
            // torsions, then impropers (dnei14_PB order); class defaults: torsions
            // by their middle atoms (SCALED_TERM default), impropers max(n_matrix)
            if (d.ntors < 0 || d.nimp < 0 || d.n14 < 0) return JP_EINVAL;
            if ((d.ntors && (!d.tors_atoms || !d.tors_fc || !d.tors_rot || !d.tors_phase || !d.tors_div)) ||
                (d.nimp && (!d.imp_atoms || !d.imp_fc || !d.imp_rot || !d.imp_phase || !d.imp_div)) ||
                (d.n14 && (!d.p14_atoms || !d.p14_vdw || !d.p14_elec)))
                return JP_EINVAL;
            for (int q = 0; q < d.ntors + d.nimp; ++q) {
                const bool imp = q >= d.ntors;
                const int u = imp ? q - d.ntors : q;
                const int32_t *at = imp ? d.imp_atoms + 4 * u : d.tors_atoms + 4 * u;
                int a4[4];
                for (int k = 0; k < 4; ++k) {
                    a4[k] = at[k] - 1;
                    if (a4[k] < 0 || a4[k] >= n) return JP_EINVAL;
                }
                const double rot = imp ? d.imp_rot[u] : d.tors_rot[u];
                const int divn = imp ? d.imp_div[u] : d.tors_div[u];
                const double nr = std::round(rot);
                if (!(nr >= 1.0 && nr <= 6.0) || divn == 0) return JP_EINVAL;   // dnetor_PB handles n = 1..6
                const int32_t *cl = imp ? d.imp_cls : d.tors_cls;
                const int c = cl ? cl[u] - 1
                                 : (imp ? cmax : d.n_matrix[(d.icls[a4[2]] - 1) * d.nfrag + (d.icls[a4[1]] - 1)] - 1);
                if (c < 0 || c >= d.nfrg) return JP_EINVAL;
                const int t = (int)tor.size();
                tor.push_back(make_int4(a4[0] + atom0, a4[1] + atom0, a4[2] + atom0, a4[3] + atom0));
                torp.push_back(make_double4(imp ? d.imp_fc[u] : d.tors_fc[u], rot, imp ? d.imp_phase[u] : d.tors_phase[u],
                                            (double)divn));
                torc.push_back(c | (imp ? 256 : 0));
                em_sys[2].push_back(s); em_cls[2].push_back(c); em_slot[2].push_back(imp ? 3 : 2);
                for (int k = 0; k < 4; ++k) bref_of[a4[k] + atom0].push_back(bref_code(t, JP_BK_TORS, k));
            }
            for (int q = 0; q < d.n14; ++q) {
                const int i = d.p14_atoms[2 * q] - 1, j = d.p14_atoms[2 * q + 1] - 1;
                if (i < 0 || j < 0 || i >= n || j >= n || i == j) return JP_EINVAL;
                const int c = d.p14_cls ? d.p14_cls[q] - 1 : d.n_matrix[(d.icls[j] - 1) * d.nfrag + (d.icls[i] - 1)] - 1;
                if (c < 0 || c >= d.nfrg) return JP_EINVAL;
                const int t = (int)p14.size();
                p14.push_back(make_int4(i + atom0, j + atom0, c, 0));
                p14p.push_back(make_double2(d.p14_vdw[q], d.p14_elec[q]));
                for (int e = 0; e < 2; ++e) { em_sys[3].push_back(s); em_cls[3].push_back(c); em_slot[3].push_back(4 + e); }
                bref_of[i + atom0].push_back(bref_code(t, JP_BK_P14, 0));
                bref_of[j + atom0].push_back(bref_code(t, JP_BK_P14, 1));
            }
        }
        for (int i = 0; i < npad; ++i) slot_sys_h.push_back(s);
        for (int i = 0; i < npad / JP_CL; ++i) clu_sys_h.push_back(s);
        atom0 += n;
        slot0 += npad;
        lj0 += nt * nt;
    }
    natom = atom0; nslot = slot0; nclu = slot0 / JP_CL;
    // bonded gather lists: atoms with terms and their codes; energy entries in
    // k_bonded_energy order (bonds, angles, torsions, 1-4 pairs x 2)
    nbt = (int)bnd.size(); nat = (int)ang.size(); ntt = (int)tor.size(); n14 = (int)p14.size();
    for (int k = 0; k < 4; ++k) {
        esys_h.insert(esys_h.end(), em_sys[k].begin(), em_sys[k].end());
        ecls_h.insert(ecls_h.end(), em_cls[k].begin(), em_cls[k].end());
        eslot_h.insert(eslot_h.end(), em_slot[k].begin(), em_slot[k].end());
    }
    std::vector<int> bat, bptr{0}, bref;
    bref_of.resize(natom);
    for (int a = 0; a < natom; ++a) {
        if (bref_of[a].empty()) continue;
        bat.push_back(a);
        bref.insert(bref.end(), bref_of[a].begin(), bref_of[a].end());
        bptr.push_back((int)bref.size());
    }
    nbat = (int)bat.size();
    // k_step units in atom order: each SHAKE group once (at its first atom), free atoms alone
    std::vector<int4> units;
    {
        std::vector<int> owner(natom, -1);   // unit code per group atom
        std::vector<int4> gu;
        for (int g = 0; g < (int)s2a.size(); ++g) gu.push_back(make_int4(s2a[g].x, s2a[g].y, -1, 2 | (g << 3)));
        for (int g = 0; g < (int)s3a.size(); ++g) gu.push_back(make_int4(s3a[g].x, s3a[g].y, s3a[g].z, 3 | (g << 3)));
        for (int g = 0; g < (int)s4a.size(); ++g) gu.push_back(make_int4(s4a[g].x, s4a[g].y, s4a[g].z, 4 | (g << 3)));
        auto unit_atom = [&](const int4 &q, int k) {
            return k == 0 ? q.x : k == 1 ? q.y : k == 2 ? q.z : s4a[q.w >> 3].w;
        };
        for (int q = 0; q < (int)gu.size(); ++q) {
            const int m[4] = {gu[q].x, gu[q].y, gu[q].z, (gu[q].w & 7) == 4 ? unit_atom(gu[q], 3) : -1};
            for (int k = 0; k < (gu[q].w & 7); ++k) {
                if (owner[m[k]] >= 0) return JP_EINVAL;   // an atom in two SHAKE groups (shake.f90:71-73)
                owner[m[k]] = q;
            }
        }
        for (int a = 0; a < natom; ++a) {
            if (owner[a] < 0) {
                units.push_back(make_int4(a, -1, -1, 1));
            } else {
                const int4 q = gu[owner[a]];
                int first = q.x;
                for (int k = 1; k < (q.w & 7); ++k) first = std::min(first, unit_atom(q, k));
                if (first == a) units.push_back(q);
            }
        }
        nunit = (int)units.size();
        units_h = units;
    }
    n2 = (int)s2a.size(); n3 = (int)s3a.size(); n4 = (int)s4a.size();
    mass_h = mass;
    set_factors(-1);
    // upload
    AL(d_sp, ns); AL(d_atom_sys, natom); AL(d_slot_sys, nslot);
    AL(d_clu_sys, nclu);
    AL(d_x, 3 * (size_t)natom); AL(d_v, 3 * (size_t)natom); AL(d_xold, 3 * (size_t)natom);
    AL(d_vback, 3 * (size_t)natom); AL(d_mass, natom); AL(d_imass, natom); AL(d_chg, natom);
    AL(d_grad, 3 * (size_t)natom); AL(d_eexc, 3 * (size_t)natom); AL(d_ekin, natom);
    AL(d_tflag, natom); AL(d_mol, natom); AL(d_xptr, natom + 1); AL(d_xidx, xidx.size());
    AL(d_eptr, natom + 1); AL(d_eidx, eidx.size()); AL(d_efac, 3 * (size_t)ns);
    AL(d_lj64, lj64.size()); AL(d_lj32, lj32.size());
    AL(d_img, 3 * (size_t)natom); AL(d_cellid, natom); AL(d_subkey, natom); AL(d_refcell, 3 * (size_t)natom);
    AL(d_tmp, natom);
    AL(d_inum, natom); AL(d_perm, nslot);
    AL(d_ys, nslot); AL(d_qd, nslot); AL(d_xr, nslot); AL(d_imgs, nslot); AL(d_lp, nslot); AL(d_mcell, nslot);
    AL(d_facc, 4 * (size_t)nslot); AL(d_icnt, nslot); AL(d_jcnt, nslot);
    AL(d_mcnt, nmcnt); AL(d_nbflag, 1);
    AL(d_wmask, wmask_h.size());
    AL(d_cbb, 6 * (size_t)nclu); AL(d_ccen, nclu); AL(d_ccf, nclu); AL(d_lcnt, nclu);
    AL(d_sysext, ns); AL(d_err, 1); AL(d_lst, 2 + (size_t)ns);
    AL(d_gslot, 3 * (size_t)nslot);
    AL(d_ep64, 6 * (size_t)nclu); AL(d_epfix, 6 * (size_t)nclu);
    AL(d_s2sys, n2); AL(d_s3sys, n3); AL(d_s2a, n2); AL(d_s3a, n3); AL(d_s2d, 6 * (size_t)n2); AL(d_s3d, 6 * (size_t)n3);
    AL(d_s4sys, n4); AL(d_s4a, n4); AL(d_s4d, 6 * (size_t)n4);
    AL(d_u64, 3);
    AL(d_bat, bat.size()); AL(d_bptr, bptr.size()); AL(d_bref, bref.size());
    AL(d_bnd, bnd.size()); AL(d_bndp, bndp.size()); AL(d_ang, ang.size()); AL(d_angp, angp.size());
    AL(d_eterm, (size_t)nbt + nat + ntt + 2 * (size_t)n14);
    AL(d_tor, tor.size()); AL(d_torp, torp.size()); AL(d_torc, torc.size()); AL(d_p14, p14.size()); AL(d_p14p, p14p.size());
    AL(d_unit, units.size());
    AL(d_abeg, ns); AL(d_aend, ns); AL(d_cbeg, ns); AL(d_cend, ns);
    AL(d_rsum, 10 * (size_t)ns); AL(d_rfix, 6 * (size_t)ns);
    AL(d_part, 8 * (size_t)JP_SEG_NB * ns); AL(d_partf, 8 * (size_t)JP_SEG_NB * ns);
    CK(hipHostMalloc((void **)&h_rsum, sizeof(double) * 10 * (size_t)ns, hipHostMallocDefault));
    CK(hipHostMalloc((void **)&h_rfix, sizeof(long long) * 6 * (size_t)ns, hipHostMallocDefault));
    CK(hipHostMalloc((void **)&h_err, 2 * sizeof(int), hipHostMallocDefault));
    CK(hipHostMalloc((void **)&h_lst, sizeof(int) * (2 + (size_t)ns), hipHostMallocDefault));
    AL(d_xbak, 3 * (size_t)natom); AL(d_vbak, 3 * (size_t)natom); AL(d_errbak, 1);
    std::vector<int> abeg(ns), aend(ns), cbeg(ns), cend(ns);
    for (int s = 0; s < ns; ++s) {
        abeg[s] = sp[s].atom0; aend[s] = sp[s].atom0 + sp[s].natom;
        cbeg[s] = sp[s].slot0 / JP_CL; cend[s] = (sp[s].slot0 + sp[s].nslot) / JP_CL;
    }
    auto H2D = [&](void *dst, const void *src, size_t bytes) -> int {
        if (bytes) CK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
        return 0;
    };
    int r = 0;
    r |= grid_tables();
    r |= upload_sp();
    r |= H2D(d_atom_sys.p, atom_sys_h.data(), 4 * atom_sys_h.size());
    r |= H2D(d_slot_sys.p, slot_sys_h.data(), 4 * slot_sys_h.size());
    r |= H2D(d_clu_sys.p, clu_sys_h.data(), 4 * clu_sys_h.size());
    r |= H2D(d_x.p, x.data(), 8 * x.size());
    r |= H2D(d_mass.p, mass.data(), 8 * mass.size());
    r |= H2D(d_imass.p, imass.data(), 8 * imass.size());
    r |= H2D(d_chg.p, chg.data(), 8 * chg.size());
    r |= H2D(d_tflag.p, tflag.data(), 4 * tflag.size());
    r |= H2D(d_mol.p, mol.data(), 4 * mol.size());
    r |= H2D(d_xptr.p, xptr.data(), 4 * xptr.size());
    r |= H2D(d_xidx.p, xidx.data(), 4 * xidx.size());
    r |= H2D(d_eptr.p, eptr.data(), 4 * eptr.size());
    r |= H2D(d_eidx.p, eidx.data(), 8 * eidx.size());
    r |= H2D(d_lj64.p, lj64.data(), 16 * lj64.size());
    r |= H2D(d_lj32.p, lj32.data(), 8 * lj32.size());
    r |= H2D(d_s2sys.p, s2sys.data(), 4 * s2sys.size());
    r |= H2D(d_s3sys.p, s3sys.data(), 4 * s3sys.size());
    r |= H2D(d_s2a.p, s2a.data(), 16 * s2a.size());
    r |= H2D(d_s3a.p, s3a.data(), 16 * s3a.size());
    r |= H2D(d_s2d.p, s2d.data(), 8 * s2d.size());
    r |= H2D(d_s3d.p, s3d.data(), 8 * s3d.size());
    r |= H2D(d_s4sys.p, s4sys.data(), 4 * s4sys.size());
    r |= H2D(d_s4a.p, s4a.data(), 16 * s4a.size());
    r |= H2D(d_s4d.p, s4d.data(), 8 * s4d.size());
    r |= H2D(d_bat.p, bat.data(), 4 * bat.size());
    r |= H2D(d_bptr.p, bptr.data(), 4 * bptr.size());
    r |= H2D(d_bref.p, bref.data(), 4 * bref.size());
    r |= H2D(d_bnd.p, bnd.data(), 16 * bnd.size());
    r |= H2D(d_bndp.p, bndp.data(), 16 * bndp.size());
    r |= H2D(d_ang.p, ang.data(), 16 * ang.size());
    r |= H2D(d_angp.p, angp.data(), 16 * angp.size());
    r |= H2D(d_tor.p, tor.data(), 16 * tor.size());
    r |= H2D(d_torp.p, torp.data(), 32 * torp.size());
    r |= H2D(d_torc.p, torc.data(), 4 * torc.size());
    r |= H2D(d_p14.p, p14.data(), 16 * p14.size());
    r |= H2D(d_p14p.p, p14p.data(), 16 * p14p.size());
    r |= H2D(d_unit.p, units.data(), 16 * units.size());
    r |= H2D(d_abeg.p, abeg.data(), 4 * abeg.size());
    r |= H2D(d_aend.p, aend.data(), 4 * aend.size());
    r |= H2D(d_cbeg.p, cbeg.data(), 4 * cbeg.size());
    r |= H2D(d_cend.p, cend.data(), 4 * cend.size());
    r |= H2D(d_wmask.p, wmask_h.data(), 8 * wmask_h.size());
    if (r) return r;
    CK(hipMemsetAsync(d_v.p, 0, 8 * 3 * (size_t)natom, st));
    CK(hipMemsetAsync(d_facc.p, 0, 32 * (size_t)nslot, st));
    CK(hipStreamSynchronize(st));
    list_half = want_half(JP_PREC_FP32_FIXED);   // jp_build_list's list until a run asks for another
    return 0;
}

// Class factors: onehot < 0 -> the run's factors (LMD: lambda^2, lambda, 1 on
// classes 1..3; verlet.f90:67-91), else 1 on class `onehot` and 0 elsewhere.
int Engine::set_factors(int onehot) {
    std::vector<double> ef(3 * (size_t)nsys, 0.0);
    for (int s = 0; s < nsys; ++s) {
        double f[3];
        for (int k = 0; k < 3; ++k) {
            if (onehot >= 0) f[k] = (k == onehot) ? 1.0 : 0.0;
            else if (hs[s].lmd && hs[s].nfrg == 3) f[k] = (k == 0) ? hs[s].lambda * hs[s].lambda : (k == 1 ? hs[s].lambda : 1.0);
            else f[k] = 1.0;
        }
        for (int q = 0; q < 4; ++q) sp[s].fac[q] = f[sp[s].cls[q]];
        for (int k = 0; k < 3; ++k) ef[3 * s + k] = f[k];
    }
    if (!d_sp.p) return 0;
    int r = upload_sp();
    if (r) return r;
    CK(hipMemcpyAsync(d_efac.p, ef.data(), 8 * ef.size(), hipMemcpyHostToDevice, st));
    return 0;
}

// Cell ranges of the systems' box grids and the per-cell tables: Hilbert
// rank -> cell, cell -> grid coordinates, cell -> system (the same curve
// k_bin bins on).  Allocates the per-cell device arrays.
int Engine::grid_tables() {
    std::vector<int> crank, ccoord, csys;
    int cell0 = 0;
    for (int s = 0; s < nsys; ++s) {
        SysP &P = sp[s];
        P.cell0 = cell0;
        P.ncell = P.nc[0] * P.nc[1] * P.nc[2];
        ccoord.resize((size_t)cell0 + P.ncell);
        for (int cz = 0; cz < P.nc[2]; ++cz)
            for (int cy = 0; cy < P.nc[1]; ++cy)
                for (int cx = 0; cx < P.nc[0]; ++cx) {
                    const int h = (int)hilbert3(cx, cy, cz, P.hbits);
                    crank.push_back(cell0 + h);
                    ccoord[(size_t)cell0 + h] = cx | (cy << 10) | (cz << 20);
                }
        csys.insert(csys.end(), (size_t)P.ncell, s);
        cell0 += P.ncell;
    }
    ncell = cell0;
    AL(d_cell_sys, ncell); AL(d_crank, ncell); AL(d_ccoord, ncell);
    AL(d_cnt, ncell + 1); AL(d_start, ncell + 1); AL(d_fill, ncell); AL(d_cslot, ncell);
    AL(d_cellext, ncell); AL(d_crec, ncell);
    CK(hipMemcpyAsync(d_cell_sys.p, csys.data(), 4 * csys.size(), hipMemcpyHostToDevice, st));
    CK(hipMemcpyAsync(d_crank.p, crank.data(), 4 * crank.size(), hipMemcpyHostToDevice, st));
    CK(hipMemcpyAsync(d_ccoord.p, ccoord.data(), 4 * ccoord.size(), hipMemcpyHostToDevice, st));
    // the host vectors go out of scope: the copies must be done
    CK(hipStreamSynchronize(st));
    cell_sys_h.swap(csys);
    return 0;
}

// A different box grid (JP_OPT_GRID_SHIFT, the autotuner): new cell tables,
// and the next evaluation rebuilds the list on it.  The pair set is the same
// on every grid; the slot order, and with it the grouping of partial sums,
// follows the grid.
int Engine::regrid(int shift) {
    for (int s = 0; s < nsys; ++s) set_grid(sp[s], shift);
    grid_shift = shift;
    int r;
    if ((r = grid_tables())) return r;
    if ((r = upload_sp())) return r;
    drop_graphs();
    have_list = false;
    return 0;
}

// celprp: bin and sort the atoms, then build the j-lists into segments of
// ccap entries per i-cluster.  One status read-back per rebuild.
int Engine::build_list(bool defer) {
    if (cmm) {   // synchronous: celprp's ier at once
        CK(hipEventRecord(ev[2], st));
        const int r = cmm_celprp();
        CK(hipEventRecord(ev[3], st));
        CK(hipEventSynchronize(ev[3]));
        float ms = 0;
        hipEventElapsedTime(&ms, ev[2], ev[3]);
        last_build_ms = ms;
        t_list += ms * 1e-3;
        n_list++;
        have_list = r == 0;
        return r;
    }
    CK(hipEventRecord(ev[2], st));
    const int nstat = 2 + nsys;
    const int nrs = std::max(std::max(std::max(ncell + 1, nslot), std::max(nsys, nstat)), nmcnt);
    hipLaunchKernelGGL(k_list_reset, grid1(nrs, 256), 256, 0, st, ncell, d_cnt.p, d_fill.p, d_cellext.p, nsys,
                       d_sysext.p, nstat, d_lst.p, nslot, d_perm.p, d_jcnt.p, nmcnt, d_mcnt.p);
    hipLaunchKernelGGL(k_bin, grid1(natom, 256), 256, 0, st, natom, d_atom_sys.p, d_sp.p, d_x.p, d_img.p,
                       d_cellid.p, d_subkey.p, d_refcell.p, d_cnt.p, d_lst.p, d_tflag.p, d_mcnt.p);
    hipLaunchKernelGGL(k_scan, 1, 1024, 0, st, ncell, d_cnt.p, d_start.p);
    hipLaunchKernelGGL(k_place, grid1(natom, 256), 256, 0, st, natom, d_cellid.p, d_subkey.p, d_start.p, d_fill.p,
                       d_tmp.p);
    hipLaunchKernelGGL(k_rank, grid1(natom, 256), 256, 0, st, natom, d_atom_sys.p, d_sp.p, d_cellid.p, d_subkey.p,
                       d_start.p, d_cnt.p, d_tmp.p, d_inum.p, d_perm.p);
    hipLaunchKernelGGL(k_cslot, grid1(ncell, 256), 256, 0, st, ncell, d_cell_sys.p, d_sp.p, d_start.p, d_cslot.p);
    hipLaunchKernelGGL(k_fill_sorted, grid1(nslot, 256), 256, 0, st, nslot, d_slot_sys.p, d_sp.p, d_perm.p, d_x.p,
                       d_img.p, d_chg.p, d_tflag.p, d_mol.p, d_ys.p, d_qd.p, d_xr.p, d_imgs.p, d_lp.p,
                       d_refcell.p, d_mcell.p);
    hipLaunchKernelGGL(k_cbox, grid1(JP_CL * nclu, 256), 256, 0, st, nclu, d_clu_sys.p, d_sp.p, d_perm.p, d_cellid.p,
                       d_ccoord.p, d_ys.p, d_lp.p, d_cbb.p, d_ccen.p, d_ccf.p, d_cellext.p);
    hipLaunchKernelGGL(k_crec, grid1(ncell, 256), 256, 0, st, ncell, d_crank.p, d_cslot.p, d_cnt.p, d_cellext.p,
                       d_crec.p, d_cell_sys.p, d_sysext.p);
    // Segments of ccap entries per i-cluster.  ccap only grows (lists above
    // 7/8 of it report their length), so captured step graphs stay valid.
    cell_err = 0;
    ccap = ccap_force > 0 ? ccap_force : std::max(ccap, ccap_want);
    ccap_force = 0;
    AL(d_ent, (size_t)nclu * ccap);
    AL(d_mb, (size_t)nclu * ccap / 8);
    launch_list8(*this, ccap - ccap / 8);
    CK(hipGetLastError());
    CK(hipMemcpyAsync(h_lst, d_lst.p, 4 * (size_t)nstat, hipMemcpyDeviceToHost, st));
    CK(hipEventRecord(ev[3], st));
    if (defer) return 0;
    return finish_list(true);
}

int Engine::finish_list(bool retry) {
    if (cmm) return 0;   // CMM rebuilds are checked when they run
    const int nstat = 2 + nsys;
    CK(hipEventSynchronize(ev[3]));
    for (int attempt = 0;; ++attempt) {
        if (h_lst[1] > 0) ccap_want = std::max(ccap_want, (h_lst[1] + h_lst[1] / 4 + 255) / 256 * 256);
        cell_err |= h_lst[0] & JP_ERR_CELL;
        if (!(h_lst[0] & JP_ERR_LISTCAP)) break;
        if (!retry) return JP_LIST_REDO;
        if (attempt >= 6 || ccap >= (1 << 20)) return JP_IER_LIST;
        ccap_want = std::max(ccap_want, 2 * ccap);
        ccap = std::max(ccap, ccap_want);
        CK(hipMemsetAsync(d_lst.p, 0, 4 * (size_t)nstat, st));
        CK(hipMemsetAsync(d_jcnt.p, 0, 4 * (size_t)nslot, st));
        AL(d_ent, (size_t)nclu * ccap);
        AL(d_mb, (size_t)nclu * ccap / 8);
        launch_list8(*this, ccap - ccap / 8);
        CK(hipGetLastError());
        CK(hipMemcpyAsync(h_lst, d_lst.p, 4 * (size_t)nstat, hipMemcpyDeviceToHost, st));
        CK(hipEventRecord(ev[3], st));
        CK(hipEventSynchronize(ev[3]));
    }
    float ms = 0;
    hipEventElapsedTime(&ms, ev[2], ev[3]);
    last_build_ms = ms;
    t_list += ms * 1e-3;
    n_list++;
    have_list = true;
    int ier = 0;
    if ((h_lst[0] | cell_err) & JP_ERR_CELL) ier = JP_IER_CELL;
    else if ((h_lst[0] & JP_ERR_NBLIM) && nblim_exceeded()) ier = JP_IER_LIST;   // cutoff.f90:25-46
    else
        for (int s = 0; s < nsys; ++s)
            if (h_lst[2 + s] != (list_half ? 1 : 2) * hs[s].nexcl) ier = JP_IER_EXCL;  // cutoff.f90:90-96
    if (ier && !retry) return JP_LIST_REDO;
    return ier;
}

// Some atom has more than n15mx listed partners: md_run's own test counts
// them per energy class and table (k_nbexact).  A half list keeps no
// per-atom partner list to count, so there the total decides.
bool Engine::nblim_exceeded() {
    if (list_half) return true;
    if (hipMemsetAsync(d_nbflag.p, 0, 4, st) != hipSuccess) return true;
    hipLaunchKernelGGL(k_nbexact, nclu, 64, 0, st, nclu, d_clu_sys.p, d_sp.p, d_ys.p, ccap, d_lcnt.p, d_ent.p,
                       d_nbflag.p);
    int f = 1;
    if (hipMemcpyAsync(&f, d_nbflag.p, 4, hipMemcpyDeviceToHost, st) != hipSuccess) return true;
    if (hipStreamSynchronize(st) != hipSuccess) return true;
    return f != 0;
}

int Engine::refresh_sorted() {
    hipLaunchKernelGGL(k_post, grid1(natom, 256), 256, 0, st, natom, d_atom_sys.p, d_sp.p, d_inum.p, d_img.p, d_x.p,
                       d_xold.p, d_v.p, d_ys.p, 0);
    if (prec_built >= 0) {
        // qd does not change between rebuilds; nothing else to refresh
    }
    CK(hipGetLastError());
    return 0;
}

template <bool E, bool M, bool S>
static void launch_pair(int prec, unsigned nb, size_t lds32, size_t lds64, hipStream_t st, Engine &g);
template <bool E, bool M, bool S>
static void launch_h32(unsigned nb, size_t lds32, hipStream_t st, Engine &g);

// Pair-kernel variant: bit 1 class factors differ from 1 (LMD passes), bit 0
// some box edge < 2*rlim (per-pair minimum image).
// reps force-step pair launches back to back between two events (one warm-up
// launch first); the pair kernel only rewrites the gradient slots.
int Engine::time_pair(int prec, int reps, double &sec) {
    int r;
    if ((r = ensure_list(prec))) return r;
    if ((r = set_factors(-1))) return r;
    if ((r = refresh_sorted())) return r;
    if ((r = nb(prec, false, false))) return r;
    CK(hipEventRecord(ev[0], st));
    for (int k = 0; k < reps; ++k)
        if ((r = nb(prec, false, false))) return r;
    CK(hipEventRecord(ev[1], st));
    if ((r = facc_zero())) return r;   // the timing launches' sums are not used
    CK(hipEventSynchronize(ev[1]));
    float ms = 0;
    CK(hipEventElapsedTime(&ms, ev[0], ev[1]));
    sec = ms * 1e-3 / reps;
    return 0;
}

int Engine::pair_variant() const {
    bool lmd = false, small = false;
    for (int s = 0; s < nsys; ++s) {
        small = small || sp[s].small;
        for (int q = 0; q < 4; ++q) lmd = lmd || sp[s].fac[q] != 1.0 || sp[s].cls[q] != sp[s].cls[0];
    }
    return (lmd ? 2 : 0) | (small ? 1 : 0) | (pw << 4) | (pwh << 8) | (list_half ? 1 << 12 : 0);
}

int Engine::pair(int prec, bool energy) {
    const unsigned nb = (unsigned)((nclu + 7) / 8 * 8);
    int maxt = 0;
    for (int s = 0; s < nsys; ++s) maxt = std::max(maxt, hs[s].ntyp);
    const size_t l32 = sizeof(float2) * maxt * maxt, l64 = sizeof(double2) * maxt * maxt;
    const int key = (energy ? 4 : 0) | (pair_variant() & 3);
    if (prec == JP_PREC_FP32_FIXED && list_half) {
        switch (key) {
            case 0: launch_h32<false, false, false>(nb, l32, st, *this); break;
            case 1: launch_h32<false, false, true>(nb, l32, st, *this); break;
            case 2: launch_h32<false, true, false>(nb, l32, st, *this); break;
            case 3: launch_h32<false, true, true>(nb, l32, st, *this); break;
            case 4: launch_h32<true, false, false>(nb, l32, st, *this); break;
            case 5: launch_h32<true, false, true>(nb, l32, st, *this); break;
            case 6: launch_h32<true, true, false>(nb, l32, st, *this); break;
            default: launch_h32<true, true, true>(nb, l32, st, *this); break;
        }
        CK(hipGetLastError());
        return 0;
    }
    switch (key) {
        case 0: launch_pair<false, false, false>(prec, nb, l32, l64, st, *this); break;
        case 1: launch_pair<false, false, true>(prec, nb, l32, l64, st, *this); break;
        case 2: launch_pair<false, true, false>(prec, nb, l32, l64, st, *this); break;
        case 3: launch_pair<false, true, true>(prec, nb, l32, l64, st, *this); break;
        case 4: launch_pair<true, false, false>(prec, nb, l32, l64, st, *this); break;
        case 5: launch_pair<true, false, true>(prec, nb, l32, l64, st, *this); break;
        case 6: launch_pair<true, true, false>(prec, nb, l32, l64, st, *this); break;
        default: launch_pair<true, true, true>(prec, nb, l32, l64, st, *this); break;
    }
    CK(hipGetLastError());
    return 0;
}

int Engine::pair_out(int prec) {
    if (prec != JP_PREC_FP32_FIXED || !list_half) return 0;
    hipLaunchKernelGGL(k_facc_out, grid1(nslot, 256), 256, 0, st, nslot, d_slot_sys.p, d_sp.p, excess_args(), d_facc.p,
                       d_gslot.p);
    CK(hipGetLastError());
    return 0;
}

int Engine::facc_zero() {
    hipLaunchKernelGGL(k_facc_zero, grid1(nslot, 256), 256, 0, st, nslot, d_facc.p);
    CK(hipGetLastError());
    return 0;
}

// ALSD: the class factors of the current lambda, before the forces.
int Engine::alsd_fac() {
    if (!alsd) return 0;
    hipLaunchKernelGGL(k_alsd_fac, grid1(nsys, 64), 64, 0, st, nsys, (const double *)d_lam.p, d_sp.p, d_efac.p);
    CK(hipGetLastError());
    return 0;
}

int Engine::nb(int prec, bool energy, bool itwin) {
    if (cmm) return cmm_forces(energy, itwin);
    int r = pair(prec, energy);
    if (!r) r = pair_out(prec);
    return r;
}

int Engine::forces(int prec, bool energy) {
    int r = alsd_fac();
    if (r) return r;
    if ((r = nb(prec, energy || ge || alsd, cmm_far_now))) return r;
    if (nbat) {
        hipLaunchKernelGGL(k_bonded, grid1(nbat, 128), 128, 0, st, bonded_args(), d_atom_sys.p, d_sp.p, d_x.p,
                           d_efac.p, d_inum.p, d_gslot.p, (double *)nullptr);
        CK(hipGetLastError());
    }
    if (alsd) return alsd_force(prec);
    return ge ? ge_weight(prec) : 0;
}

// ALSD: the step's class-1 and class-2 energies from device sums (as
// ge_weight), then lambda's force and half-step velocity (k_alsd).
int Engine::alsd_force(int prec) {
    int r;
    if (cmm) {
        if ((r = cmm_sums())) return r;
        prec = JP_PREC_FP64;
    } else {
        hipLaunchKernelGGL(k_excess_energy, grid1(natom, 256), 256, 0, st, natom, d_atom_sys.p, d_sp.p, d_x.p,
                           d_chg.p, d_eptr.p, d_eidx.p, d_eexc.p);
        if ((r = seg_sum<double>(d_abeg.p, d_aend.p, 3, d_eexc.p, d_part.p, d_rsum.p + 6 * (size_t)nsys))) return r;
        if (prec == JP_PREC_FP64) r = seg_sum<double>(d_cbeg.p, d_cend.p, 6, d_ep64.p, d_part.p, d_rsum.p);
        else r = seg_sum<long long>(d_cbeg.p, d_cend.p, 6, d_epfix.p, d_partf.p, d_rfix.p);
        if (r) return r;
    }
    const int ne = nbt + nat + ntt + 2 * n14;
    if (ne) {
        hipLaunchKernelGGL(k_bonded_energy, grid1(ne, 256), 256, 0, st, nbt, nat, ntt, n14, bonded_args(),
                           d_atom_sys.p, d_sp.p, d_x.p, d_eterm.p);
        CK(hipGetLastError());
    }
    AlsdArgs A;
    A.nw = al_nw; A.nd = al_nd; A.par = d_alpar.p;
    A.rsum = d_rsum.p;
    A.rfix = prec == JP_PREC_FP64 ? nullptr : d_rfix.p;
    A.nsys = nsys;
    A.hf = (prec == JP_PREC_FP32_FIXED && list_half) ? 1.0 : 0.5;
    A.dself = d_dself.p;
    A.eterm = d_eterm.p; A.esys = d_esys.p; A.ecls = d_ecls.p; A.eslot = d_eslot.p; A.ne = ne;
    A.lmass = d_lmass.p; A.lam = d_lam.p; A.sp = d_sp.p;
    hipLaunchKernelGGL(k_alsd, nsys, 256, 0, st, A);
    CK(hipGetLastError());
    return 0;
}

int Engine::lam_upload() {
    std::vector<double> L((size_t)JP_LS * nsys, 0.0);
    for (int s = 0; s < nsys; ++s) {
        L[(size_t)JP_LS * s] = hs[s].lambda;
        L[(size_t)JP_LS * s + 1] = hs[s].lambda_v;
        L[(size_t)JP_LS * s + 4] = hs[s].lambda;
        L[(size_t)JP_LS * s + 5] = hs[s].lambda_v;
    }
    AL(d_lam, L.size());
    CK(hipMemcpyAsync(d_lam.p, L.data(), 8 * L.size(), hipMemcpyHostToDevice, st));
    CK(hipStreamSynchronize(st));
    return 0;
}

int Engine::lam_download() {
    std::vector<double> L((size_t)JP_LS * nsys);
    CK(hipMemcpyAsync(L.data(), d_lam.p, 8 * L.size(), hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    for (int s = 0; s < nsys; ++s) {
        hs[s].lambda = L[(size_t)JP_LS * s];
        hs[s].lambda_v = L[(size_t)JP_LS * s + 1];
    }
    return 0;
}

// ALSD for the following runs (after jp_set_canonical; three energy
// classes, GEPSME= LMD): the parameters of the NAMEGE= file (inmcmd.f90),
// lambda's mass (lmass <= 0: the class-1 atoms' mass added in atom order,
// inpcls.f90:348-357) and l_temp_control.  lambda's degree of freedom goes
// to the one temperature group, or under temperature groups to the first
// with the most degrees of freedom (md.f90:114-123, 221-228).
int Engine::set_alsd(int on, int nw, int nd, const double *c, const double *low, const double *high, double lower,
                     double upper, double celw, double ceup, double tempce, double lmass, int tctl) {
    drop_graphs();
    if (alsd) {
        int r = lam_download();
        if (r) return r;
    }
    alsd = false;
    if (cano && !gdof_h.empty()) {   // degrees of freedom without lambda's
        CK(hipMemcpyAsync(d_gdof.p, gdof_h.data(), 8 * gdof_h.size(), hipMemcpyHostToDevice, st));
        CK(hipStreamSynchronize(st));
    }
    if (!on) {
        set_factors(-1);
        return 0;
    }
    if (!cano) return JP_EINVAL;
    for (int s = 0; s < nsys; ++s)
        if (!hs[s].lmd || hs[s].nfrg != 3) return JP_EINVAL;
    if (nw < 1 || nd < 0 || nw > 4096 || nd > 64 || !c || !low || !high) return JP_EINVAL;
    std::vector<double> par{lower, upper, celw, ceup, tempce};
    for (int w = 0; w < nw; ++w) par.push_back(low[w]);
    for (int w = 0; w < nw; ++w) par.push_back(high[w]);
    for (size_t q = 0; q < (size_t)nw * (nd + 1); ++q) par.push_back(c[q]);
    for (double p : par)
        if (!std::isfinite(p)) return JP_EINVAL;
    // lambda's mass per system
    lmass_h.assign(nsys, 0.0);
    for (int s = 0; s < nsys; ++s) {
        if (lmass > 0.0) {
            lmass_h[s] = lmass;
        } else {
            double m = 0.0;
            for (int a = sp[s].atom0; a < sp[s].atom0 + hs[s].n; ++a)
                if (atom_cls_h[a] == 0) m = m + mass_h[a];
            lmass_h[s] = m;
        }
        if (!(lmass_h[s] > 0.0) || !std::isfinite(lmass_h[s])) return JP_EINVAL;
    }
    // lambda's temperature group and degree of freedom
    std::vector<double> dof = gdof_h;
    lgrp_h.assign(nsys, -1);
    for (int s = 0; s < nsys; ++s) {
        if (!cano_grouped) {
            lgrp_h[s] = sg0_h[s];   // caltmp always adds lambda's kinetic term
            if (tctl) dof[sg0_h[s]] += 1.0;
        } else if (tctl) {
            int g = sg0_h[s];
            for (int q = sg0_h[s] + 1; q < sg1_h[s]; ++q)
                if (dof[q] > dof[g]) g = q;
            lgrp_h[s] = g;
            dof[g] += 1.0;
        }
    }
    const int ne = nbt + nat + ntt + 2 * n14;
    std::vector<double> ds(3 * (size_t)nsys);
    for (int s = 0; s < nsys; ++s)
        for (int k = 0; k < 3; ++k) ds[3 * s + k] = hs[s].dself[k];
    AL(d_alpar, par.size());
    AL(d_lmass, nsys); AL(d_lgrp, nsys); AL(d_dself, 3 * (size_t)nsys);
    AL(d_eslot, std::max(ne, 1)); AL(d_esys, std::max(ne, 1)); AL(d_ecls, std::max(ne, 1));
    AL(d_lam, (size_t)JP_LS * nsys); AL(d_lamb, (size_t)JP_LS * nsys); AL(d_lamb2, (size_t)JP_LS * nsys);
    CK(hipMemcpyAsync(d_alpar.p, par.data(), 8 * par.size(), hipMemcpyHostToDevice, st));
    CK(hipMemcpyAsync(d_lmass.p, lmass_h.data(), 8 * (size_t)nsys, hipMemcpyHostToDevice, st));
    CK(hipMemcpyAsync(d_lgrp.p, lgrp_h.data(), 4 * (size_t)nsys, hipMemcpyHostToDevice, st));
    CK(hipMemcpyAsync(d_dself.p, ds.data(), 8 * ds.size(), hipMemcpyHostToDevice, st));
    CK(hipMemcpyAsync(d_gdof.p, dof.data(), 8 * dof.size(), hipMemcpyHostToDevice, st));
    if (ne) {
        CK(hipMemcpyAsync(d_eslot.p, eslot_h.data(), 4 * (size_t)ne, hipMemcpyHostToDevice, st));
        CK(hipMemcpyAsync(d_esys.p, esys_h.data(), 4 * (size_t)ne, hipMemcpyHostToDevice, st));
        CK(hipMemcpyAsync(d_ecls.p, ecls_h.data(), 4 * (size_t)ne, hipMemcpyHostToDevice, st));
    }
    CK(hipStreamSynchronize(st));
    al_nw = nw;
    al_nd = nd;
    al_tctl = tctl ? 1 : 0;
    ge = false;
    alsd = true;
    return lam_upload();
}

// METHOD= GE: the step's class-1 potential energy from device sums (excess
// energies per atom, pair energies per cluster -- the pair kernel ran in
// energy mode --, bonded energies per term), then its multicanonical weight
// into d_gew (jp_ge.hpp).  Launches only: capturable in a step graph.
int Engine::ge_weight(int prec) {
    int r;
    if (cmm) {
        if ((r = cmm_sums())) return r;
        prec = JP_PREC_FP64;
    } else {
        hipLaunchKernelGGL(k_excess_energy, grid1(natom, 256), 256, 0, st, natom, d_atom_sys.p, d_sp.p, d_x.p,
                           d_chg.p, d_eptr.p, d_eidx.p, d_eexc.p);
        if ((r = seg_sum<double>(d_abeg.p, d_aend.p, 3, d_eexc.p, d_part.p, d_rsum.p + 6 * (size_t)nsys))) return r;
        if (prec == JP_PREC_FP64) r = seg_sum<double>(d_cbeg.p, d_cend.p, 6, d_ep64.p, d_part.p, d_rsum.p);
        else r = seg_sum<long long>(d_cbeg.p, d_cend.p, 6, d_epfix.p, d_partf.p, d_rfix.p);
        if (r) return r;
    }
    const int ne = nbt + nat + ntt + 2 * n14;
    if (ne) {
        hipLaunchKernelGGL(k_bonded_energy, grid1(ne, 256), 256, 0, st, nbt, nat, ntt, n14, bonded_args(),
                           d_atom_sys.p, d_sp.p, d_x.p, d_eterm.p);
        CK(hipGetLastError());
    }
    GeArgs G;
    G.nw = ge_nw; G.nd = ge_nd; G.par = d_gepar.p;
    G.rsum = d_rsum.p;
    G.rfix = prec == JP_PREC_FP64 ? nullptr : d_rfix.p;
    G.nsys = nsys;
    G.hf = (prec == JP_PREC_FP32_FIXED && list_half) ? 1.0 : 0.5;
    G.dself = hs[0].dself[0];
    G.eterm = d_eterm.p; G.eslot = d_eslot.p; G.ne = ne;
    G.out = d_gew.p;
    hipLaunchKernelGGL(k_ge_weight, 1, 256, 0, st, G);
    CK(hipGetLastError());
    return 0;
}

// METHOD= GE for the following runs (with METHOD= CANO's temperature control
// set by jp_set_canonical): one system, one energy class.
int Engine::set_mcmd(int on, int nw, int nd, const double *c, const double *low, const double *high,
                     double alphalw, double alphaup, double celw, double ceup, double tempce) {
    drop_graphs();
    ge = false;
    if (!on) return 0;
    if (alsd) {
        int r = set_alsd(0, 0, 0, nullptr, nullptr, nullptr, 0, 0, 0, 0, 0, 0, 0);
        if (r) return r;
    }
    if (nsys != 1 || hs[0].nfrg != 1 || hs[0].lmd) return JP_EINVAL;
    if (nw < 1 || nd < 0 || nw > 4096 || nd > 64 || !c || !low || !high) return JP_EINVAL;
    if (!(celw < ceup) || !std::isfinite(alphalw) || !std::isfinite(alphaup) || !std::isfinite(tempce))
        return JP_EINVAL;
    std::vector<double> par;
    par.push_back(alphalw); par.push_back(alphaup); par.push_back(celw); par.push_back(ceup); par.push_back(tempce);
    for (int w = 0; w < nw; ++w) par.push_back(low[w]);
    for (int w = 0; w < nw; ++w) par.push_back(high[w]);
    for (size_t q = 0; q < (size_t)nw * (nd + 1); ++q) par.push_back(c[q]);
    for (double p : par)
        if (!std::isfinite(p)) return JP_EINVAL;
    AL(d_gepar, par.size());
    AL(d_gew, 2);
    const int ne = nbt + nat + ntt + 2 * n14;
    AL(d_eslot, std::max(ne, 1));
    CK(hipMemcpyAsync(d_gepar.p, par.data(), 8 * par.size(), hipMemcpyHostToDevice, st));
    if (ne) CK(hipMemcpyAsync(d_eslot.p, eslot_h.data(), 4 * (size_t)ne, hipMemcpyHostToDevice, st));
    CK(hipStreamSynchronize(st));
    ge_nw = nw;
    ge_nd = nd;
    ge = true;
    return 0;
}

// Per-system, per-class bonded energies at the device coordinates, summed
// per energy slot in term order (dnergv_PB's routines): et[(s*3 + class)*6 + slot].
int Engine::bonded_energies(std::vector<double> &et) {
    et.assign(18 * (size_t)nsys, 0.0);
    const int ne = nbt + nat + ntt + 2 * n14;
    if (!ne) return 0;
    hipLaunchKernelGGL(k_bonded_energy, grid1(ne, 256), 256, 0, st, nbt, nat, ntt, n14, bonded_args(), d_atom_sys.p,
                       d_sp.p, d_x.p, d_eterm.p);
    CK(hipGetLastError());
    std::vector<double> e(ne);
    CK(hipMemcpyAsync(e.data(), d_eterm.p, 8 * (size_t)ne, hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    for (int q = 0; q < ne; ++q) {
        double &dst = et[(3 * (size_t)esys_h[q] + ecls_h[q]) * 6 + eslot_h[q]];
        dst = dst + e[q];
    }
    return 0;
}

// dnergv_PB's bonded part for jp_bonded: per-class gradient accumulated into
// grad(3,N,nfrg), per-class energies eterm(6,nfrg).
int Engine::bonded(const double *cord, double *grad, double *eterm) {
    int r;
    if (cord) CK(hipMemcpyAsync(d_x.p, cord, 24 * (size_t)natom, hipMemcpyHostToDevice, st));
    const int nfrg = hs[0].nfrg;
    if (grad && nbat) {
        std::vector<double> g(3 * (size_t)natom);
        for (int k = 0; k < nfrg; ++k) {
            if ((r = set_factors(k))) return r;
            CK(hipMemsetAsync(d_grad.p, 0, 8 * g.size(), st));
            hipLaunchKernelGGL(k_bonded, grid1(nbat, 128), 128, 0, st, bonded_args(), d_atom_sys.p, d_sp.p, d_x.p,
                               d_efac.p, d_inum.p, (double *)nullptr, d_grad.p);
            CK(hipGetLastError());
            CK(hipMemcpyAsync(g.data(), d_grad.p, 8 * g.size(), hipMemcpyDeviceToHost, st));
            CK(hipStreamSynchronize(st));
            double *gk = grad + (size_t)k * 3 * natom;
            for (size_t q = 0; q < g.size(); ++q) gk[q] += g[q];
        }
        if ((r = set_factors(-1))) return r;
    }
    std::vector<double> et;
    if ((r = bonded_energies(et))) return r;
    if (eterm)
        for (int k = 0; k < nfrg; ++k)
            for (int q = 0; q < 6; ++q) eterm[6 * k + q] = et[6 * k + q];
    CK(hipStreamSynchronize(st));
    return 0;
}

template <bool E, bool M, bool S, int PW>
static void launch_h32_pw(unsigned nb, size_t lds32, hipStream_t st, Engine &g) {
    hipLaunchKernelGGL((k_pair_h32<E, M, S, PW>), nb, 64 * PW, lds32, st, g.nclu, g.d_clu_sys.p, g.d_sp.p, g.d_ccen.p,
                       g.d_ys.p, g.d_lj32.p, g.ccap, g.d_lcnt.p, g.d_ent.p, g.d_facc.p, g.d_epfix.p);
}

// Half-list FP32 kernel: force steps at the tuned wave count, energy steps
// at JP_PWH.
template <bool E, bool M, bool S>
static void launch_h32(unsigned nb, size_t lds32, hipStream_t st, Engine &g) {
    switch (E ? JP_PWH : g.pwh) {
        case 1: launch_h32_pw<E, M, S, 1>(nb, lds32, st, g); break;
        case 4: launch_h32_pw<E, M, S, 4>(nb, lds32, st, g); break;
        case 8: launch_h32_pw<E, M, S, 8>(nb, lds32, st, g); break;
        default: launch_h32_pw<E, M, S, 2>(nb, lds32, st, g); break;
    }
}

template <bool E, bool M, bool S, int PW>
static void launch_pair_f32_only(unsigned nb, size_t lds32, hipStream_t st, Engine &g) {
    hipLaunchKernelGGL((k_pair_f32<E, M, S, PW>), nb, 64 * PW, lds32, st, g.nclu, g.d_clu_sys.p, g.d_sp.p, g.d_ccen.p,
                       g.d_ys.p, g.d_lj32.p, g.ccap, g.d_lcnt.p, g.d_ent.p, g.d_mb.p, g.excess_args(), g.d_gslot.p,
                       g.d_epfix.p);
}

template <bool E, bool M, bool S, int PW>
static void launch_pair_pw(int prec, unsigned nb, size_t lds32, size_t lds64, hipStream_t st, Engine &g) {
    if (prec == JP_PREC_FP64)
        hipLaunchKernelGGL((k_pair_f64<E, M, S, PW>), nb, 64 * PW, lds64, st, g.nclu, g.d_clu_sys.p, g.d_sp.p,
                           g.d_ccen.p, g.d_ys.p, g.d_qd.p, g.d_lj64.p, g.ccap, g.d_lcnt.p, g.d_ent.p, g.excess_args(),
                           g.d_gslot.p, g.d_ep64.p);
    else
        hipLaunchKernelGGL((k_pair_f32<E, M, S, PW>), nb, 64 * PW, lds32, st, g.nclu, g.d_clu_sys.p, g.d_sp.p,
                           g.d_ccen.p, g.d_ys.p, g.d_lj32.p, g.ccap, g.d_lcnt.p, g.d_ent.p, g.d_mb.p, g.excess_args(),
                           g.d_gslot.p, g.d_epfix.p);
}

// The tuned wave count applies to the FP32 force-step kernels; energy steps
// and the plain FP64 path keep JP_PW.
template <bool E, bool M, bool S>
static void launch_pair(int prec, unsigned nb, size_t lds32, size_t lds64, hipStream_t st, Engine &g) {
    if constexpr (E) {
        launch_pair_pw<E, M, S, JP_PW>(prec, nb, lds32, lds64, st, g);
    } else {
        if (prec == JP_PREC_FP64 || g.pw == JP_PW) launch_pair_pw<E, M, S, JP_PW>(prec, nb, lds32, lds64, st, g);
        else if (g.pw == 2) launch_pair_f32_only<E, M, S, 2>(nb, lds32, st, g);
        else launch_pair_f32_only<E, M, S, 8>(nb, lds32, st, g);
    }
}

template <int LW, bool HALF, int NB>
static void launch_list8_lw(Engine &g, int hiwater) {
    hipLaunchKernelGGL((k_list8<LW, HALF, NB>), grid1(g.nclu, LW), 64 * LW, 0, g.st, g.nclu, g.d_clu_sys.p, g.d_sp.p,
                       g.d_cbb.p, g.d_ccf.p, g.d_crec.p, g.d_sysext.p, g.d_lp.p, g.d_xr.p, g.d_imgs.p, g.d_xptr.p,
                       g.d_xidx.p, g.d_mcell.p, g.cell_table ? g.d_wmask.p : nullptr, g.ccap, hiwater, g.d_ent.p,
                       reinterpret_cast<uint8_t *>(g.d_mb.p), g.d_lcnt.p, g.d_lst.p, g.d_icnt.p, g.d_jcnt.p);
    if (HALF)
        hipLaunchKernelGGL(k_nbcheck, grid1(g.nslot, 256), 256, 0, g.st, g.nslot, g.d_slot_sys.p, g.d_sp.p, g.d_icnt.p,
                           g.d_jcnt.p, g.d_lst.p);
}

template <bool HALF, int NB>
static void launch_list8_h(Engine &g, int hiwater) {
    switch (g.lw) {
        case 1: launch_list8_lw<1, HALF, NB>(g, hiwater); break;
        case 2: launch_list8_lw<2, HALF, NB>(g, hiwater); break;
        case 8: launch_list8_lw<8, HALF, NB>(g, hiwater); break;
        default: launch_list8_lw<4, HALF, NB>(g, hiwater); break;
    }
}

// The half list keeps the default batch count.
static void launch_list8(Engine &g, int hiwater) {
    if (g.list_half) launch_list8_h<true, JP_LNB>(g, hiwater);
    else if (g.lnb == 1) launch_list8_h<false, 1>(g, hiwater);
    else if (g.lnb == 4) launch_list8_h<false, 4>(g, hiwater);
    else launch_list8_h<false, JP_LNB>(g, hiwater);
}

// Autotuner (where SIMD_check.f90:2-42 doubles i_vec while the pair routine
// gets faster).  Per box-grid cell size (the 4.5-9 A grid, and grids with
// cells half and twice as wide): the list kernel at 1, 2, 4 and 8
// wavefronts per block and 1, 2 and 4 batches of 8 j-clusters per pass (the
// tile width), the FP32 force-step pair kernel at 2, 4 and 8 wavefronts per
// i-cluster, each on the live list (median of 7 launches, candidates
// interleaved so that clock drift hits all of them alike); then whole
// rebuilds at the best list shape.  The grid with the least rebuild time /
// update + pair time wins, with its shapes, and the list is rebuilt on it.
// The pair set is the same on every grid and shape; the grid and pw regroup
// the partial sums (deterministic for given choices, within the FP32 bar
// across them, and FP64 sums within round-off).
int Engine::tune(int prec, int update) {
    int r;
    hipEvent_t a, b;
    CK(hipEventCreate(&a));
    CK(hipEventCreate(&b));
    struct EvGuard {
        hipEvent_t &a, &b;
        ~EvGuard() { hipEventDestroy(a); hipEventDestroy(b); }
    } guard{a, b};
    constexpr int NREP = 7;
    auto once = [&](auto launch) -> float {
        float ms = 0.f;
        hipEventRecord(a, st);
        launch();
        hipEventRecord(b, st);
        hipEventSynchronize(b);
        hipEventElapsedTime(&ms, a, b);
        return ms;
    };
    auto median = [](std::vector<float> t) {
        std::sort(t.begin(), t.end());
        return t[t.size() / 2];
    };
    // fastest candidate of a set of knob settings; returns its median time
    auto pick = [&](const std::vector<std::pair<int, int>> &cand, int &k1, int &k2, auto launch) -> float {
        std::vector<std::vector<float>> t(cand.size());
        for (int rep = 0; rep < NREP; ++rep)
            for (size_t q = 0; q < cand.size(); ++q) {
                k1 = cand[q].first;
                k2 = cand[q].second;
                t[q].push_back(once(launch));
            }
        float best = 1e30f;
        size_t bq = 0;
        for (size_t q = 0; q < cand.size(); ++q) {
            const float m = median(t[q]);
            if (m < best) { best = m; bq = q; }
        }
        k1 = cand[bq].first;
        k2 = cand[bq].second;
        return best;
    };
    int dummy = 0;
    std::vector<std::pair<int, int>> lcand, pcand;
    for (int nb : {1, 2, 4})
        for (int w : {1, 2, 4, 8}) lcand.push_back({w, nb});
    if (prec == JP_PREC_FP64) pcand = {{JP_PW, 0}};
    else if (list_half) pcand = {{1, 0}, {2, 0}, {4, 0}, {8, 0}};
    else pcand = {{2, 0}, {4, 0}, {8, 0}};
    const int shift0 = grid_shift;
    struct Choice { int shift, lw, lnb, pw; float score; };
    Choice best{shift0, lw, lnb, pw, 1e30f};
    std::vector<std::vector<int>> seen;
    for (int dsh : {0, -1, 1}) {
        const int shift = shift0 + dsh;
        // skip a shift that gives every system a grid already timed
        std::vector<int> hb;
        for (int s = 0; s < nsys; ++s) {
            SysP q = sp[s];
            set_grid(q, shift);
            hb.push_back(q.hbits);
        }
        if (std::find(seen.begin(), seen.end(), hb) != seen.end()) continue;
        seen.push_back(hb);
        if (shift != grid_shift && (r = regrid(shift))) return r;
        if ((r = ensure_list(prec))) return r;
        pick(lcand, lw, lnb, [&] { launch_list8(*this, ccap - ccap / 8); });
        int &pk = list_half ? pwh : pw;
        float tp;
        if (prec == JP_PREC_FP64) {
            std::vector<float> t;
            for (int rep = 0; rep < NREP; ++rep) t.push_back(once([&] { pair(prec, false); }));
            tp = median(t);
        } else {
            tp = pick(pcand, pk, dummy, [&] { pair(prec, false); });
        }
        // whole rebuilds at this shape (set-up kernels, list, status read-back)
        std::vector<float> tb;
        for (int rep = 0; rep < 3; ++rep) {
            if ((r = build_list())) return r;
            tb.push_back(last_build_ms);
        }
        const float score = median(tb) / (float)std::max(1, update) + tp;
        if (score < best.score) best = Choice{shift, lw, lnb, list_half ? pwh : pw, score};
    }
    if (best.shift != grid_shift && (r = regrid(best.shift))) return r;
    lw = best.lw;
    lnb = best.lnb;
    if (prec != JP_PREC_FP64) (list_half ? pwh : pw) = best.pw;
    CK(hipGetLastError());
    int r2 = facc_zero();
    if (r2) return r2;
    tuned = 1;
    return build_list();
}

// Per-system, per-class energies after a pair(prec, true) call.
int Engine::energies(int prec, std::vector<double> &eel, std::vector<double> &evd, std::vector<double> &eex) {
    eel.assign(3 * (size_t)nsys, 0.0);
    evd.assign(3 * (size_t)nsys, 0.0);
    eex.assign(3 * (size_t)nsys, 0.0);
    // per-atom excess energies and per-cluster pair energies, summed per
    // system on the device in a fixed order; one small read-back.  CMM: the
    // per-atom near and far energies, no excess or self term.
    int r;
    if (cmm) {
        if ((r = cmm_sums())) return r;
        prec = JP_PREC_FP64;
    } else {
        hipLaunchKernelGGL(k_excess_energy, grid1(natom, 256), 256, 0, st, natom, d_atom_sys.p, d_sp.p, d_x.p,
                           d_chg.p, d_eptr.p, d_eidx.p, d_eexc.p);
        if ((r = seg_sum<double>(d_abeg.p, d_aend.p, 3, d_eexc.p, d_part.p, d_rsum.p + 6 * (size_t)nsys))) return r;
        if (prec == JP_PREC_FP64) r = seg_sum<double>(d_cbeg.p, d_cend.p, 6, d_ep64.p, d_part.p, d_rsum.p);
        else r = seg_sum<long long>(d_cbeg.p, d_cend.p, 6, d_epfix.p, d_partf.p, d_rfix.p);
        if (r) return r;
    }
    CK(hipMemcpyAsync(h_rsum, d_rsum.p, sizeof(double) * 9 * (size_t)nsys, hipMemcpyDeviceToHost, st));
    if (prec != JP_PREC_FP64)
        CK(hipMemcpyAsync(h_rfix, d_rfix.p, sizeof(long long) * 6 * (size_t)nsys, hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    for (int s = 0; s < nsys; ++s)
        for (int k = 0; k < 3; ++k) {
            if (prec == JP_PREC_FP64) {
                eel[3 * s + k] = h_rsum[6 * s + k];
                evd[3 * s + k] = h_rsum[6 * s + 3 + k];
            } else {
                eel[3 * s + k] = h_rfix[6 * s + k] / JP_ESCALE;
                evd[3 * s + k] = h_rfix[6 * s + 3 + k] / JP_ESCALE;
            }
            eex[3 * s + k] = h_rsum[6 * (size_t)nsys + 3 * s + k];
        }
    // the full list counts every pair from both sides (enear.f90:568-571), the half list once
    const double hf = (prec == JP_PREC_FP32_FIXED && list_half) ? 1.0 : 0.5;
    for (int s = 0; s < nsys; ++s) {
        for (int k = 0; k < 3; ++k) {
            eel[3 * s + k] = eel[3 * s + k] * hf + hs[s].dself[k] + eex[3 * s + k];
            evd[3 * s + k] = evd[3 * s + k] * hf;
        }
    }
    return 0;
}

int Engine::nonbond(const double *cord, double *grad, double *eelec, double *evdw, int prec) {
    int r;
    if (cord) CK(hipMemcpyAsync(d_x.p, cord, 24 * (size_t)natom, hipMemcpyHostToDevice, st));
    if (!have_list || list_half != want_half(prec)) {
        list_half = want_half(prec);
        if ((r = build_list())) return r;
    } else if ((r = refresh_sorted())) {
        return r;
    }
    int nfrg = hs[0].nfrg;
    std::vector<double> eel, evd, eex;
    std::vector<double> g(3 * (size_t)natom);
    for (int k = 0; k < nfrg; ++k) {
        if ((r = set_factors(nfrg == 1 ? -1 : k))) return r;
        // CMM: the far field once, at the first class (md_run's itwin = 1)
        if ((r = nb(prec, k == 0, k == 0))) return r;
        if (k == 0 && (r = energies(prec, eel, evd, eex))) return r;
        hipLaunchKernelGGL(k_kick, grid1(natom, 256), 256, 0, st, natom, d_atom_sys.p, d_sp.p, d_inum.p, d_gslot.p,
                           d_mass.p, d_x.p, d_v.p, d_xold.p, (double *)nullptr, d_grad.p, 0);
        CK(hipGetLastError());
        if (grad) {
            CK(hipMemcpyAsync(g.data(), d_grad.p, 8 * g.size(), hipMemcpyDeviceToHost, st));
            CK(hipStreamSynchronize(st));
            double *gk = grad + (size_t)k * 3 * natom;
            for (size_t q = 0; q < g.size(); ++q) gk[q] += g[q];
        }
    }
    if ((r = set_factors(-1))) return r;
    for (int k = 0; k < nfrg; ++k) {
        if (eelec) eelec[k] = eel[k];
        if (evdw) evdw[k] = evd[k];
    }
    CK(hipStreamSynchronize(st));
    return 0;
}

int Engine::shake(int &ier_out) {
    ier_out = 0;
    if (n2) hipLaunchKernelGGL(k_shake2, grid1(n2, 128), 128, 0, st, n2, d_s2sys.p, d_sp.p, d_s2a.p, d_s2d.p, d_mass.p,
                               d_x.p, d_xold.p);
    if (n3) hipLaunchKernelGGL(k_shake3, grid1(n3, 128), 128, 0, st, n3, d_s3sys.p, d_sp.p, d_s3a.p, d_s3d.p, d_imass.p,
                               d_x.p, d_xold.p, shake_tol, shake_maxit, d_err.p);
    if (n4) hipLaunchKernelGGL(k_shake4, grid1(n4, 128), 128, 0, st, n4, d_s4sys.p, d_sp.p, d_s4a.p, d_s4d.p,
                               d_imass.p, d_x.p, d_xold.p, shake_tol, shake_maxit, d_err.p);
    CK(hipGetLastError());
    return 0;
}

// k_step over every unit: with 4-atom SHAKE groups in the run the variant
// that holds four atoms per unit (shake4_PB), else three.
template <bool KICK, bool VUPD, int NQ>
static void launch_step_nq(Engine &g, double *vb) {
    hipLaunchKernelGGL((k_step<KICK, VUPD, NQ>), grid1(g.nunit, 128), 128, 0, g.st, g.nunit, g.d_unit.p,
                       g.d_atom_sys.p, g.d_sp.p, g.d_inum.p, g.d_img.p, g.d_gslot.p, g.d_mass.p, g.d_imass.p, g.d_s2d.p,
                       g.d_s3d.p, g.d_s4a.p, g.d_s4d.p, g.shake_tol, g.shake_maxit, (g.n2 + g.n3 + g.n4) > 0 ? 1 : 0,
                       g.d_x.p, g.d_v.p, vb, g.d_ys.p, g.d_err.p);
}

template <bool KICK, bool VUPD>
static void launch_step(Engine &g, double *vb) {
    if (g.n4) launch_step_nq<KICK, VUPD, 4>(g, vb);
    else launch_step_nq<KICK, VUPD, 3>(g, vb);
}

int Engine::integrate(int prec, bool mon) {
    (void)prec;
    if (cano) return cano_step();
    double *vb = mon ? d_vback.p : (double *)nullptr;
    if (stop_mode) {   // verlet.f90:144-190: kick, STOPCE about x(t), drift and SHAKE
        hipLaunchKernelGGL(k_kick, grid1(natom, 256), 256, 0, st, natom, d_atom_sys.p, d_sp.p, d_inum.p, d_gslot.p,
                           d_mass.p, d_imass.p, 0, d_v.p, vb);
        hipLaunchKernelGGL(k_remotv, (unsigned)nsys, JP_STOP_NT, 0, st, d_stop.p, d_mass.p, d_x.p, d_v.p,
                           (const int *)nullptr, 1, d_err.p);
        launch_step<false, true>(*this, vb);
    } else {
        launch_step<true, true>(*this, vb);
    }
    CK(hipGetLastError());
    return 0;
}

// STOPCE= TRAN (1), ROTA (2), BOTH (3) or NO (0) for every system (remotv,
// verlet.f90:156-169, verlethe.f90:489-502, 705-718, staver.f90:165-178):
// the first natom[s] atoms of system s are stopped (md_run's
// ixcend(NSTPCN)); nkind[s] atoms from its first share atom 1's molecule
// kind (md.f90:187-193, the degrees of freedom under TEMPCO= YES).  Before
// jp_set_canonical, which takes the degrees of freedom off.
int Engine::set_stopce(int mode, const int *natom, const int *nkind) {
    if (mode < 0 || mode > 3 || cano) return JP_EINVAL;
    std::vector<int4> h(nsys, make_int4(0, 0, 0, 0));
    if (mode) {
        if (!natom) return JP_EINVAL;
        for (int s = 0; s < nsys; ++s) {
            const int k = nkind ? nkind[s] : 0;
            if (natom[s] < 1 || natom[s] > hs[s].n || k < 0 || k > hs[s].n) return JP_EINVAL;
            h[s] = make_int4(mode, sp[s].atom0, natom[s], k);
        }
    }
    drop_graphs();
    AL(d_stop, nsys);
    CK(hipMemcpyAsync(d_stop.p, h.data(), sizeof(int4) * nsys, hipMemcpyHostToDevice, st));
    CK(hipStreamSynchronize(st));
    stop_h = h;
    stop_mode = mode;
    return 0;
}

// STOPCE's degrees of freedom (md.f90:114-121, 183-215): 3 (TRAN, ROTA) or
// 6 (BOTH) off the system's group labelled 1 (its only group without
// labels), or, when the atoms of atom 1's molecule kind carry labels up to
// ifrg >= 2, off the groups labelled 1 to ifrg in proportion to their
// degrees of freedom.  keys: the sorted (system, label) of every group.
void Engine::stop_dof(std::vector<double> &dof, const std::vector<long long> &keys, const int *group) const {
    if (!stop_mode) return;
    for (int s = 0; s < nsys; ++s) {
        const int4 S = stop_h[s];
        const double k = S.x == 3 ? 6.0 : 3.0;
        int ifrg = 1;
        if (group)
            for (int a = S.y; a < S.y + S.w; ++a) ifrg = std::max(ifrg, group[a]);
        std::vector<int> gs;   // the system's groups labelled 1..ifrg, by label
        for (int g = 0; g < (int)keys.size(); ++g)
            if ((int)(keys[g] >> 32) == s && (int)(keys[g] & 0xffffffff) <= ifrg) gs.push_back(g);
        if (gs.empty())
            for (int g = 0; g < (int)keys.size(); ++g)
                if ((int)(keys[g] >> 32) == s) { gs.push_back(g); break; }
        if (gs.empty()) continue;
        if (ifrg >= 2) {
            double sum = 0.0;
            for (int g : gs) sum = sum + dof[g];
            std::vector<double> ratio;
            for (int g : gs) ratio.push_back(dof[g] / sum);
            for (size_t i = 0; i < gs.size(); ++i) dof[gs[i]] = dof[gs[i]] - k * ratio[i];
        } else {
            dof[gs[0]] = dof[gs[0]] - k;
        }
    }
}

// md_run's initial step of a run without a restart (inista, md.f90:1056-1065;
// staver.f90:27-207): the list of x, the forces at x (the lambda-combined
// gradient; no METHOD= GE weight, lambda stays), the half kick
// v + (-dt c / 2) g (1/m), STOPCE about x, the drift and SHAKE of the
// coordinates; v keeps the half-kick velocity.  idloop stays (0 for
// md_run's fresh start); under CMM the far field is computed.
int Engine::start_step(int prec) {
    int r;
    if ((r = ensure_list(prec))) return r;
    if ((r = set_factors(-1))) return r;
    if ((r = refresh_sorted())) return r;
    if (alsd && (r = lam_upload())) return r;
    CK(hipMemsetAsync(d_err.p, 0, 4, st));
    cmm_far_now = true;
    if ((r = alsd_fac()) || (r = nb(prec, false, cmm_far_now))) return r;
    if (nbat) {
        hipLaunchKernelGGL(k_bonded, grid1(nbat, 128), 128, 0, st, bonded_args(), d_atom_sys.p, d_sp.p, d_x.p,
                           d_efac.p, d_inum.p, d_gslot.p, (double *)nullptr);
    }
    hipLaunchKernelGGL(k_kick, grid1(natom, 256), 256, 0, st, natom, d_atom_sys.p, d_sp.p, d_inum.p, d_gslot.p,
                       d_mass.p, d_imass.p, 1, d_v.p, (double *)nullptr);
    if (stop_mode)
        hipLaunchKernelGGL(k_remotv, (unsigned)nsys, JP_STOP_NT, 0, st, d_stop.p, d_mass.p, d_x.p, d_v.p,
                           (const int *)nullptr, 1, d_err.p);
    launch_step<false, false>(*this, (double *)nullptr);
    CK(hipGetLastError());
    int err = 0;
    CK(hipMemcpyAsync(&err, d_err.p, 4, hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    if (alsd && (r = lam_download())) return r;
    return step_ier(err);
}

// METHOD= CANO: temperature groups (one per system, or per (system, label)
// with labels such as md_run's ixamol2), their degrees of freedom (3 per atom
// minus the SHAKE constraints of the groups whose first atom is theirs,
// md.f90:114-118, 165-172) and the round buffers.
int Engine::set_canonical(int on, double temps, int icslop, double fcstol, const int *group) {
    if (on && nsys == 1) {
        AL(d_xs, 3 * (size_t)natom); AL(d_vs, 3 * (size_t)natom); AL(d_errs, 1);
        AL(d_xs2, 3 * (size_t)natom); AL(d_vs2, 3 * (size_t)natom);
    }
    drop_graphs();
    cano = false;   // METHOD= MICR unless the new set-up completes
    if (alsd) {     // ALSD needs the canonical set-up it was given with
        int r = lam_download();
        if (r) return r;
        alsd = false;
        set_factors(-1);
    }
    if (!on) return 0;
    if (!(temps > 0.0) || icslop < 1 || !(fcstol >= 0.0)) return JP_EINVAL;
    std::vector<long long> key(natom);
    for (int a = 0; a < natom; ++a) {
        const int lab = group ? group[a] : 1;
        if (lab < 1) return JP_EINVAL;
        key[a] = ((long long)atom_sys_h[a] << 32) | lab;
    }
    std::vector<long long> keys(key);
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    ngrp = (int)keys.size();
    std::vector<int> gid(natom), gsys(ngrp);
    std::vector<double> dof(ngrp, 0.0);
    for (int a = 0; a < natom; ++a) {
        gid[a] = (int)(std::lower_bound(keys.begin(), keys.end(), key[a]) - keys.begin());
        dof[gid[a]] += 3.0;
    }
    for (int g = 0; g < ngrp; ++g) gsys[g] = (int)(keys[g] >> 32);
    std::vector<int> shkon(nsys, 0), sg0(nsys, ngrp), sg1(nsys, 0), sfirst(nsys);
    // blocks of the canonical kernels: up to 256 atoms or 128 units of one
    // system, counted from the system's first (units follow the atom order)
    std::vector<int2> ablk, ublk;
    std::vector<int> ab0(nsys), ab1(nsys), ub0(nsys), ub1(nsys), ubeg(nsys + 1, nunit);
    for (int u = nunit - 1; u >= 0; --u) ubeg[atom_sys_h[units_h[u].x]] = u;
    for (int s = nsys - 1; s >= 0; --s) ubeg[s] = std::min(ubeg[s], ubeg[s + 1]);
    for (int s = 0; s < nsys; ++s) {
        ab0[s] = (int)ablk.size();
        for (int a = sp[s].atom0; a < sp[s].atom0 + hs[s].n; a += 256)
            ablk.push_back(make_int2(a, std::min(a + 256, sp[s].atom0 + hs[s].n)));
        ab1[s] = (int)ablk.size();
        ub0[s] = (int)ublk.size();
        for (int u = ubeg[s]; u < ubeg[s + 1]; u += 128) ublk.push_back(make_int2(u, std::min(u + 128, ubeg[s + 1])));
        ub1[s] = (int)ublk.size();
    }
    nab = (int)ablk.size();
    nub = (int)ublk.size();
    for (const int4 &U : units_h) {
        const int kind = U.w & 7;
        if (kind < 2) continue;
        dof[gid[U.x]] -= (double)(kind * (kind - 1) / 2);
        shkon[atom_sys_h[U.x]] = 1;
    }
    stop_dof(dof, keys, group);
    for (int g = 0; g < ngrp; ++g) {
        if (!(dof[g] > 0.0)) return JP_EINVAL;   // md.f90:124-129
        const int s = (int)(keys[g] >> 32);
        sg0[s] = std::min(sg0[s], g);
        sg1[s] = std::max(sg1[s], g + 1);
    }
    for (int s = 0; s < nsys; ++s) sfirst[s] = sp[s].atom0;
    // a block reads its system's group temperatures into JP_CANO_MAXG slots
    for (int s = 0; s < nsys; ++s)
        if (sg1[s] - sg0[s] > JP_CANO_MAXG) return JP_EINVAL;
    std::vector<double> tset(nsys, temps);
    AL(d_gid, natom); AL(d_gsys, ngrp); AL(d_gdof, ngrp);
    AL(d_sg0, nsys); AL(d_sg1, nsys); AL(d_ab0, nsys); AL(d_ab1, nsys); AL(d_ub0, nsys); AL(d_ub1, nsys);
    AL(d_sfirst, nsys); AL(d_shkon, nsys); AL(d_cdone, nsys); AL(d_ctemps, nsys); AL(d_ttgt, nsys);
    AL(d_fcc, 3 * (size_t)natom); AL(d_xcb, 3 * (size_t)natom);
    AL(d_pa, (size_t)ngrp * nab); AL(d_pu, 2 * (size_t)ngrp * nub); AL(d_pn, 2 * (size_t)ngrp * nub); AL(d_ablk, nab); AL(d_ublk, nub);
    auto H2D = [&](void *dst, const void *src, size_t bytes) -> int {
        if (bytes) CK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
        return 0;
    };
    int r = 0;
    r |= H2D(d_gid.p, gid.data(), 4 * (size_t)natom);
    r |= H2D(d_gsys.p, gsys.data(), 4 * (size_t)ngrp);
    r |= H2D(d_ab0.p, ab0.data(), 4 * (size_t)nsys);
    r |= H2D(d_ab1.p, ab1.data(), 4 * (size_t)nsys);
    r |= H2D(d_ub0.p, ub0.data(), 4 * (size_t)nsys);
    r |= H2D(d_ub1.p, ub1.data(), 4 * (size_t)nsys);
    r |= H2D(d_ablk.p, ablk.data(), 8 * (size_t)nab);
    r |= H2D(d_ublk.p, ublk.data(), 8 * (size_t)nub);
    r |= H2D(d_gdof.p, dof.data(), 8 * (size_t)ngrp);
    r |= H2D(d_sg0.p, sg0.data(), 4 * (size_t)nsys);
    r |= H2D(d_sg1.p, sg1.data(), 4 * (size_t)nsys);
    r |= H2D(d_sfirst.p, sfirst.data(), 4 * (size_t)nsys);
    r |= H2D(d_shkon.p, shkon.data(), 4 * (size_t)nsys);
    r |= H2D(d_ctemps.p, tset.data(), 8 * (size_t)nsys);
    r |= H2D(d_ttgt.p, tset.data(), 8 * (size_t)nsys);
    if (r) return r;
    CK(hipStreamSynchronize(st));   // the host vectors end with this call
    cano_temps = temps;
    ttgt_set = temps;
    cano_icslop = icslop;
    cano_fcstol = fcstol;
    cano_grouped = group != nullptr;
    gdof_h = dof;
    sg0_h = sg0;
    sg1_h = sg1;
    cano = true;
    return 0;
}

__global__ void k_fill_d(double *p, int n, double v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// The canonical target temperature of the steps that follow, in stream
// order (a fill kernel: the heating steps are host launches).
int Engine::put_target(double t) {
    if (t == ttgt_set) return 0;
    hipLaunchKernelGGL(k_fill_d, grid1(nsys, 64), 64, 0, st, d_ttgt.p, nsys, t);
    CK(hipGetLastError());
    ttgt_set = t;
    return 0;
}

int Engine::set_heating(int htlp, double tmpsta) {
    if (htlp < 0 || !std::isfinite(tmpsta) || tmpsta < 0.0) return JP_EINVAL;
    heat_lp = htlp;
    heat_t0 = tmpsta;
    return 0;
}

CanoArgs Engine::cano_args() {
    CanoArgs A;
    A.natom = natom; A.nunit = nunit; A.icslop = cano_icslop; A.nab = nab; A.nub = nub; A.ngrp = ngrp;
    A.fcstol = cano_fcstol;
    A.atom_sys = d_atom_sys.p; A.inum = d_inum.p; A.img = d_img.p;
    A.gid = d_gid.p; A.gsys = d_gsys.p; A.sg0 = d_sg0.p; A.sg1 = d_sg1.p;
    A.ab0 = d_ab0.p; A.ab1 = d_ab1.p; A.ub0 = d_ub0.p; A.ub1 = d_ub1.p; A.sfirst = d_sfirst.p;
    A.ablk = d_ablk.p; A.ublk = d_ublk.p;
    A.shk_on = d_shkon.p; A.gdof = d_gdof.p; A.temps = d_ctemps.p; A.ttgt = d_ttgt.p;
    A.mass = d_mass.p; A.imass = d_imass.p; A.gslot = d_gslot.p;
    A.gew = ge ? d_gew.p : nullptr;
    A.lam = alsd ? d_lam.p : nullptr;
    A.lmass = d_lmass.p;
    A.lgrp = d_lgrp.p;
    A.ltc = al_tctl;
    A.x = d_x.p; A.v = d_v.p; A.xold = d_xold.p; A.vback = d_vback.p; A.fcc = d_fcc.p; A.xcb = d_xcb.p;
    A.pa = d_pa.p; A.pu = d_pu.p; A.pn = d_pn.p; A.done = d_cdone.p;
    return A;
}

// One METHOD= CANO step after the forces (jp_cano.hpp): 1 + icslop launches
// (2 without SHAKE groups); the rounds of a system that has ended are skipped
// on the device, and there is none after the last.
// One launch of a SHAKE/temperature round (phase PH of k_cano_shake), with
// four atoms per unit when the run has 4-atom SHAKE groups.
template <int PH>
static void launch_cano_round(Engine &g, const CanoArgs &A, int round) {
    if (g.n4)
        hipLaunchKernelGGL((k_cano_shake<PH, 4>), (unsigned)g.nub, 128, 0, g.st, A, g.d_sp.p, g.d_unit.p, g.d_s2d.p,
                           g.d_s3d.p, g.d_s4a.p, g.d_s4d.p, g.shake_tol, g.shake_maxit, round, g.d_ys.p, g.d_err.p);
    else
        hipLaunchKernelGGL((k_cano_shake<PH, 3>), (unsigned)g.nub, 128, 0, g.st, A, g.d_sp.p, g.d_unit.p, g.d_s2d.p,
                           g.d_s3d.p, g.d_s4a.p, g.d_s4d.p, g.shake_tol, g.shake_maxit, round, g.d_ys.p, g.d_err.p);
}

int Engine::cano_step() {
    const CanoArgs A = cano_args();
    const unsigned ga = (unsigned)nab;
    hipLaunchKernelGGL(k_cano_begin, ga, 256, 0, st, A, d_sp.p);
    if (n2 + n3 + n4 > 0) {
        for (int round = 1; round <= cano_icslop; ++round) {
            if (!stop_mode) {
                launch_cano_round<0>(*this, A, round);
                continue;
            }
            launch_cano_round<1>(*this, A, round);
            hipLaunchKernelGGL(k_remotv, (unsigned)nsys, JP_STOP_NT, 0, st, d_stop.p, d_mass.p, d_xold.p, d_v.p,
                               (const int *)d_cdone.p, round, d_err.p);
            launch_cano_round<2>(*this, A, round);
        }
    } else if (!stop_mode) {
        hipLaunchKernelGGL(k_cano_drift, ga, 256, 0, st, A, d_sp.p, d_ys.p, 0);
    } else {
        hipLaunchKernelGGL(k_cano_drift, ga, 256, 0, st, A, d_sp.p, d_ys.p, 1);
        hipLaunchKernelGGL(k_remotv, (unsigned)nsys, JP_STOP_NT, 0, st, d_stop.p, d_mass.p, d_xold.p, d_v.p,
                           (const int *)nullptr, 1, d_err.p);
        hipLaunchKernelGGL(k_cano_drift, ga, 256, 0, st, A, d_sp.p, d_ys.p, 2);
    }
    CK(hipGetLastError());
    return 0;
}

// x, v (and the SHAKE status) to or from a backup; under ALSD lambda's
// state goes along (one backup per x buffer).
int Engine::save_state(double *x, double *v, int *e) {
    CK(hipMemcpyAsync(x, d_x.p, 24 * (size_t)natom, hipMemcpyDeviceToDevice, st));
    CK(hipMemcpyAsync(v, d_v.p, 24 * (size_t)natom, hipMemcpyDeviceToDevice, st));
    if (e) CK(hipMemcpyAsync(e, d_err.p, 4, hipMemcpyDeviceToDevice, st));
    if (alsd)
        CK(hipMemcpyAsync(x == d_xs2.p ? d_lamb2.p : d_lamb.p, d_lam.p, 8 * (size_t)JP_LS * nsys,
                          hipMemcpyDeviceToDevice, st));
    if (cmm) {   // the far field a redone step may reuse
        double *w = x == d_xs2.p ? d_cmm_wb2.p : d_cmm_wb.p;
        CK(hipMemcpyAsync(w, d_cmm_gwk.p, 72 * (size_t)natom, hipMemcpyDeviceToDevice, st));
        CK(hipMemcpyAsync(w + 9 * (size_t)natom, d_cmm_efar.p, 24 * (size_t)natom, hipMemcpyDeviceToDevice, st));
    }
    return 0;
}

int Engine::load_state(const double *x, const double *v, const int *e) {
    CK(hipMemcpyAsync(d_x.p, x, 24 * (size_t)natom, hipMemcpyDeviceToDevice, st));
    CK(hipMemcpyAsync(d_v.p, v, 24 * (size_t)natom, hipMemcpyDeviceToDevice, st));
    if (e) CK(hipMemcpyAsync(d_err.p, e, 4, hipMemcpyDeviceToDevice, st));
    if (alsd)
        CK(hipMemcpyAsync(d_lam.p, x == d_xs2.p ? d_lamb2.p : d_lamb.p, 8 * (size_t)JP_LS * nsys,
                          hipMemcpyDeviceToDevice, st));
    if (cmm) {
        const double *w = x == d_xs2.p ? d_cmm_wb2.p : d_cmm_wb.p;
        CK(hipMemcpyAsync(d_cmm_gwk.p, w, 72 * (size_t)natom, hipMemcpyDeviceToDevice, st));
        CK(hipMemcpyAsync(d_cmm_efar.p, w + 9 * (size_t)natom, 24 * (size_t)natom, hipMemcpyDeviceToDevice, st));
    }
    return refresh_sorted();
}

int Engine::shake_failed(bool &failed) {
    CK(hipMemcpyAsync(h_err, d_err.p, 4, hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    failed = (*h_err & 12) != 0;
    return 0;
}

// One METHOD= CANO step with verlethe's SHAKE retry (verlethe.f90:552-573,
// 745-752): a substep whose SHAKE fails is undone (x and v of its start),
// the step length halves and the forces are evaluated again; after two
// successful half substeps the length doubles back, and the step ends when
// substeps adding up to dt are done.  A substep below dt/100 ends the run
// with the SHAKE ier (md_run stops there).
int Engine::cano_retry_step(int prec) {
    int r;
    const double dt0 = sp[0].dt;
    double T = dt0;
    int ishk = 0, jshk = 1;
    ++n_retry;
    // every substep evaluates the forces with the step's itwin (cmm_far_now)
    for (;;) {
        if ((r = save_state(d_xs2.p, d_vs2.p, nullptr))) break;
        if (sp[0].dt != T) {
            sp[0].dt = T;
            if ((r = upload_sp())) break;
        }
        if ((r = forces(prec, false)) || (r = cano_step())) break;
        bool failed = false;
        if ((r = shake_failed(failed))) break;
        if (failed) {
            CK(hipMemsetAsync(d_err.p, 0, 4, st));
            if ((r = load_state(d_xs2.p, d_vs2.p, nullptr))) break;
            if (T < 0.01 * dt0) { r = JP_IER_SHAKE_CONV; break; }
            jshk *= 2;
            ishk *= 2;
            T *= 0.5;
            continue;
        }
        ++ishk;
        if (ishk == jshk) break;
        if (ishk % 2 == 0) {
            jshk /= 2;
            ishk /= 2;
            T *= 2.0;
        }
    }
    sp[0].dt = dt0;
    int r2 = upload_sp();
    CK(hipStreamSynchronize(st));
    return r ? r : r2;
}

// n plain METHOD= CANO steps launched one by one, each checked for a SHAKE
// failure and, where one occurs, redone with the retry.
int Engine::cano_steps_eager(int prec, int n) {
    int r;
    for (int k = 0; k < n; ++k) {
        cmm_far_now = cmm_due(k);
        if ((r = save_state(d_xs.p, d_vs.p, d_errs.p))) return r;
        if ((r = forces(prec, false)) || (r = integrate(prec, false))) return r;
        bool failed = false;
        if ((r = shake_failed(failed))) return r;
        if (failed) {
            if ((r = load_state(d_xs.p, d_vs.p, d_errs.p))) return r;
            if ((r = cano_retry_step(prec))) return r;
        }
    }
    return 0;
}

// md_run's restart set-up (md.f90:1104-1263, inirst): one step from the
// restart state, whose energies are md_run's energy-continuity check
// (kinetic energy of the mean of the velocities before and after the step);
// then the coordinates go back to the restart's and the velocities become
// those the step's last verlethe substep started from (inirst restores vel
// from verlethe's velback, md.f90:1240-1250): the restart's own velocities
// unless a SHAKE retry split the step.  idloop does not advance.
int Engine::restart_step(int prec, jp_energy_log *log) {
    int r;
    if ((r = ensure_list(prec))) return r;
    if ((r = set_factors(-1))) return r;
    if ((r = refresh_sorted())) return r;
    const bool retry = cano_retry();
    if (alsd && (r = lam_upload())) return r;
    CK(hipMemsetAsync(d_err.p, 0, 4, st));
    CK(hipMemcpyAsync(d_xbak.p, d_x.p, 24 * (size_t)natom, hipMemcpyDeviceToDevice, st));
    cmm_far_now = cmm_itwin;
    if (cano && (r = put_target(heat_target(idloop)))) return r;   // inirst: md_run's idloop is the restart's
    if ((r = alsd_fac()) || (r = nb(prec, true, cmm_far_now))) return r;
    if (nbat) {
        hipLaunchKernelGGL(k_bonded, grid1(nbat, 128), 128, 0, st, bonded_args(), d_atom_sys.p, d_sp.p, d_x.p,
                           d_efac.p, d_inum.p, d_gslot.p, (double *)nullptr);
        CK(hipGetLastError());
    }
    if (ge && (r = ge_weight(prec))) return r;
    if (alsd && (r = alsd_force(prec))) return r;
    std::vector<double> eel, evd, eex, et;
    if ((r = energies(prec, eel, evd, eex)) || (r = bonded_energies(et))) return r;
    if (retry && (r = save_state(d_xs.p, d_vs.p, d_errs.p))) return r;
    if ((r = integrate(prec, true))) return r;
    if (retry) {
        bool failed = false;
        if ((r = shake_failed(failed))) return r;
        if (failed) {
            if ((r = load_state(d_xs.p, d_vs.p, d_errs.p))) return r;
            if ((r = cano_retry_step(prec))) return r;
        }
    }
    // one record per system: ekin from the mean of v before and after
    const int rl = log ? mon_record((int)idloop, 0, 1, 0, eel, evd, et, log) : 0;
    int err = 0;
    CK(hipMemcpyAsync(&err, d_err.p, 4, hipMemcpyDeviceToHost, st));
    CK(hipMemcpyAsync(d_x.p, d_xbak.p, 24 * (size_t)natom, hipMemcpyDeviceToDevice, st));
    CK(hipMemcpyAsync(d_v.p, d_vback.p, 24 * (size_t)natom, hipMemcpyDeviceToDevice, st));
    CK(hipMemsetAsync(d_err.p, 0, 4, st));
    CK(hipStreamSynchronize(st));
    if ((r = refresh_sorted())) return r;
    if (alsd && (r = lam_upload())) return r;   // lambda and its velocity back too (md.f90:1255-1258)
    if (rl) return rl;
    return step_ier(err);
}

void Engine::drop_graphs() {
    for (auto &g : graphs)
        if (g.exec) hipGraphExecDestroy(g.exec);
    graphs.clear();
}

// Graph of n plain steps for the current list segment size and buffers.
Engine::StepGraph *Engine::step_graph(int n, int prec) {
    const int variant = pair_variant();
    const int ckey = cmm ? 2 * (int)(idloop % cmm_upd) + (cmm_itwin ? 1 : 0) : -1;
    // the list a graph's kernels read: ZD's entries and segment size, CMM's
    // second-nearest table and its grid
    const void *lent = cmm ? (const void *)d_cmm_l2idx.p : (const void *)d_ent.p;
    const int lcap = cmm ? cmm_vcap : ccap;
    for (auto &g : graphs)
        if (g.exec && g.nsteps == n && g.prec == prec && g.ccap == lcap && g.ent == lent && g.variant == variant &&
            g.cmm_key == ckey)
            return &g;
    // the list buffers changed: stale graphs go
    for (auto it = graphs.begin(); it != graphs.end();) {
        if (it->ccap != lcap || it->ent != lent) {
            if (it->exec) hipGraphExecDestroy(it->exec);
            it = graphs.erase(it);
        } else {
            ++it;
        }
    }
    StepGraph g;
    g.nsteps = n; g.prec = prec; g.ccap = lcap; g.ent = lent; g.variant = variant; g.cmm_key = ckey;
    ++n_captures;
    hipGraph_t graph = nullptr;
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return nullptr;
    int r = 0;
    for (int k = 0; k < n && !r; ++k) {
        cmm_far_now = cmm_due(k);
        r = forces(prec, false);
        if (!r) r = integrate(prec, false);
    }
    const hipError_t ec = hipStreamEndCapture(st, &graph);
    if (r || ec != hipSuccess || !graph) {
        if (graph) hipGraphDestroy(graph);
        (void)hipGetLastError();
        return nullptr;
    }
    const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (ei != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    graphs.push_back(g);
    return &graphs.back();
}

int Engine::mon_record(int step, int nsteps, int nmon, int nlog, const std::vector<double> &eel,
                       const std::vector<double> &evd, const std::vector<double> &et, jp_energy_log *log) {
    hipLaunchKernelGGL(k_ekin, grid1(natom, 256), 256, 0, st, natom, d_mass.p, d_v.p, d_vback.p, d_ekin.p);
    if (int r = seg_sum<double>(d_abeg.p, d_aend.p, 1, d_ekin.p, d_part.p, d_rsum.p + 9 * (size_t)nsys)) return r;
    CK(hipMemcpyAsync(h_rsum + 9 * (size_t)nsys, d_rsum.p + 9 * (size_t)nsys, sizeof(double) * nsys,
                      hipMemcpyDeviceToHost, st));
    CK(hipMemcpyAsync(h_err, d_err.p, 4, hipMemcpyDeviceToHost, st));
    std::vector<double> lam;
    if (alsd) {
        lam.resize((size_t)JP_LS * nsys);
        CK(hipMemcpyAsync(lam.data(), d_lam.p, 8 * lam.size(), hipMemcpyDeviceToHost, st));
    }
    CK(hipStreamSynchronize(st));
    const int err = *h_err;
    if (int e = step_ier(err)) return e;
    for (int s = 0; s < nsys; ++s) {
        double e = h_rsum[9 * (size_t)nsys + s];
        // ALSD: caltmp starts with lambda's term, lambda_vt the mean of its
        // velocity before and after the step (md.f90:628-633)
        const double *ls = alsd ? lam.data() + (size_t)JP_LS * s : nullptr;
        if (ls) {
            const double lvt = 0.5 * (ls[5] + ls[1]);
            e = lmass_h[s] * lvt * lvt + e;
        }
        // caltmp.f90:24-43; its factor 0.0005 is a single-precision literal
        e = e * 1.0e7 / 4.184 * (double)0.0005f;
        jp_energy_log &L = log[(size_t)s * (nsteps / nmon + 1) + nlog];
        L.step = step;
        L.ekin = e;
        L.lambda = ls ? ls[0] : hs[s].lambda;
        L.lambda_v = ls ? ls[1] : 0.0;
        const double lam_e = ls ? ls[4] : hs[s].lambda;   // the lambda the step's energies were evaluated at
        double ep = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double f = (hs[s].lmd && hs[s].nfrg == 3) ? (k == 0 ? lam_e * lam_e : (k == 1 ? lam_e : 1.0)) : 1.0;
            L.eelec[k] = k < hs[s].nfrg ? eel[3 * s + k] : 0.0;
            L.evdw[k] = k < hs[s].nfrg ? evd[3 * s + k] : 0.0;
            const double *b = et.data() + (3 * (size_t)s + k) * 6;   // bond, angle, torsion, improper, 1-4 vdW/elec
            const bool on = k < hs[s].nfrg;
            L.ebond[k] = on ? b[0] : 0.0;
            L.eangle[k] = on ? b[1] : 0.0;
            L.etors[k] = on ? b[2] : 0.0;
            L.eimp[k] = on ? b[3] : 0.0;
            L.evdw14[k] = on ? b[4] : 0.0;
            L.eelec14[k] = on ? b[5] : 0.0;
            if (on) ep += f * (b[0] + b[1] + b[2] + b[3] + b[4] + b[5] + evd[3 * s + k] + eel[3 * s + k]);
        }
        L.epot = ep;
        L.etot = ep + e;
    }
    return 0;
}

// METHOD= MICR loop (md.f90:519-694): per step the forces (pair kernel with
// the excess term), kick and drift, SHAKE, v = (x - x_old)/dt; every
// `update` steps a rebuild after the step (md.f90:655-660); energies every
// `nmon` steps.  Runs of plain steps between those points go through cached
// HIP graphs.
int Engine::run_nve(int nsteps, int update, int nmon, int prec, jp_energy_log *log) {
    int r;
    if (nsteps < 0 || update <= 0 || (ge && !cano)) return JP_EINVAL;
    if ((r = ensure_list(prec))) return r;
    if ((r = set_factors(-1))) return r;
    if ((r = refresh_sorted())) return r;
    if (autotune && !tuned && !cmm && (r = tune(prec, update))) return r;
    if (alsd && (r = lam_upload())) return r;
    CK(hipMemsetAsync(d_err.p, 0, 4, st));
    t_pair = 0; t_step = 0; n_pair = 0; n_pair_steps = nsteps;
    pe_pending = false;
    const bool retry = cano_retry();
    n_retry = 0;
    const int cap0 = n_captures;
    const double tl0 = t_list;
    // records of the monitoring steps first, the rest marked unused
    if (nmon > 0 && log)
        for (size_t q = 0; q < (size_t)nsys * (nsteps / nmon + 1); ++q) log[q].step = -1;
    CK(hipEventRecord(ev[0], st));
    int nlog = 0;
    int step = 1;
    // A rebuild after a graph segment is checked while the next graph segment
    // runs: x, v and the SHAKE flags are saved first, and a failed or
    // overflowed rebuild restores them, rebuilds synchronously and replays
    // the segment (or returns md_run's ier from the state it refers to).
    bool pending = false;
    while (step <= nsteps) {
        // md_run monitors at idloop multiples of OUTLOG, idloop counting from
        // the restart's loop number (md.f90:604-633)
        const bool mon = nmon > 0 && (idloop + 1) % nmon == 0;
        // plain steps up to the next rebuild, energy step or the end
        int nplain = 0;
        if (!mon && use_graphs && !heating(idloop + 1)) {
            const int to_rebuild = (int)(update - idloop % update);
            int lim = std::min(nsteps - step + 1, to_rebuild);
            if (nmon > 0) lim = std::min(lim, (int)(nmon - (idloop + 1) % nmon));   // stop before the energy step
            nplain = lim;
        }
        if (nplain >= 2) {
            if (cano && (r = put_target(cano_temps))) return r;
            StepGraph *g = step_graph(nplain, prec);
            if (g) {
                if (retry && (r = save_state(d_xs.p, d_vs.p, d_errs.p))) return r;
                if (pending) {
                    CK(hipMemcpyAsync(d_xbak.p, d_x.p, 24 * (size_t)natom, hipMemcpyDeviceToDevice, st));
                    CK(hipMemcpyAsync(d_vbak.p, d_v.p, 24 * (size_t)natom, hipMemcpyDeviceToDevice, st));
                    CK(hipMemcpyAsync(d_errbak.p, d_err.p, 4, hipMemcpyDeviceToDevice, st));
                    if (alsd) {
                        AL(d_lambk, (size_t)JP_LS * nsys);
                        CK(hipMemcpyAsync(d_lambk.p, d_lam.p, 8 * (size_t)JP_LS * nsys, hipMemcpyDeviceToDevice,
                                          st));
                    }
                }
                CK(hipGraphLaunch(g->exec, st));
                if (pending) {
                    pending = false;
                    r = finish_list(false);
                    if (r == JP_LIST_REDO) {
                        CK(hipStreamSynchronize(st));
                        CK(hipMemcpyAsync(d_x.p, d_xbak.p, 24 * (size_t)natom, hipMemcpyDeviceToDevice, st));
                        CK(hipMemcpyAsync(d_v.p, d_vbak.p, 24 * (size_t)natom, hipMemcpyDeviceToDevice, st));
                        CK(hipMemcpyAsync(d_err.p, d_errbak.p, 4, hipMemcpyDeviceToDevice, st));
                        if (alsd)
                            CK(hipMemcpyAsync(d_lam.p, d_lambk.p, 8 * (size_t)JP_LS * nsys,
                                              hipMemcpyDeviceToDevice, st));
                        if ((r = build_list())) return r;
                        continue;   // same segment again, on the new list
                    }
                    if (r) return r;
                }
                if (retry) {   // a SHAKE failure in the segment: redo it step by step
                    bool failed = false;
                    if ((r = shake_failed(failed))) return r;
                    if (failed) {
                        if ((r = load_state(d_xs.p, d_vs.p, d_errs.p))) return r;
                        if ((r = cano_steps_eager(prec, nplain))) return r;
                    }
                }
                step += nplain;
                idloop += nplain;
                cmm_itwin = idloop % cmm_upd == 0;   // md.f90:679-684
                if (idloop % update == 0) {
                    if ((r = build_list(true))) return r;
                    pending = true;
                }
                continue;
            }
            use_graphs = 0;   // capture unavailable: plain launches from here on
        }
        if (pending) {   // nothing ran after the rebuild: check it in place
            pending = false;
            if ((r = finish_list(true))) return r;
        }
        // the pair launch of a host-launched step is timed with one pair of
        // events; energy steps synchronise anyway, so their time is added at once,
        // other steps are timed when the next one would reuse the events
        if (pe_pending) {
            CK(hipEventSynchronize(ev[5]));
            float m2 = 0;
            CK(hipEventElapsedTime(&m2, ev[4], ev[5]));
            t_pair += m2 * 1e-3;
            pe_pending = false;
        }
        if (cano && (r = put_target(heat_target(idloop + 1)))) return r;
        if ((r = alsd_fac())) return r;
        cmm_far_now = cmm_itwin;
        CK(hipEventRecord(ev[4], st));
        if ((r = nb(prec, mon || ge || alsd, cmm_far_now))) return r;
        CK(hipEventRecord(ev[5], st));
        pe_pending = true;
        if (nbat) {
            hipLaunchKernelGGL(k_bonded, grid1(nbat, 128), 128, 0, st, bonded_args(), d_atom_sys.p, d_sp.p, d_x.p,
                               d_efac.p, d_inum.p, d_gslot.p, (double *)nullptr);
            CK(hipGetLastError());
        }
        if (ge && (r = ge_weight(prec))) return r;
        if (alsd && (r = alsd_force(prec))) return r;
        ++n_pair;
        std::vector<double> eel, evd, eex, et;
        if (mon && (r = energies(prec, eel, evd, eex))) return r;
        if (mon && (r = bonded_energies(et))) return r;
        if (retry && (r = save_state(d_xs.p, d_vs.p, d_errs.p))) return r;
        if ((r = integrate(prec, mon))) return r;
        if (retry) {
            bool failed = false;
            if ((r = shake_failed(failed))) return r;
            if (failed) {
                if ((r = load_state(d_xs.p, d_vs.p, d_errs.p))) return r;
                if ((r = cano_retry_step(prec))) return r;
            }
        }
        if (mon) {
            if ((r = mon_record((int)(idloop + 1), nsteps, nmon, nlog, eel, evd, et, log))) return r;
            nlog++;
        }
        ++idloop;
        ++step;
        cmm_itwin = idloop % cmm_upd == 0;
        if (idloop % update == 0) {     // md.f90:655-660, after the step
            if ((r = build_list())) return r;
        }
    }
    if (pending && (r = finish_list(true))) return r;
    CK(hipEventRecord(ev[1], st));
    int err = 0;
    CK(hipMemcpyAsync(&err, d_err.p, 4, hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    float ms = 0;
    CK(hipEventElapsedTime(&ms, ev[0], ev[1]));
    t_total = ms * 1e-3;
    if (pe_pending) {
        float m2 = 0;
        CK(hipEventElapsedTime(&m2, ev[4], ev[5]));
        t_pair += m2 * 1e-3;
        pe_pending = false;
    }
    const double tl = t_list - tl0;
    t_step = t_total - (n_pair ? t_pair / n_pair * nsteps : 0.0) - tl;
    t_list = tl;
    captures_last_run = n_captures - cap0;
    if (alsd && (r = lam_download())) return r;
    return step_ier(err);
}

// Captures the step graphs a following run_nve(nsteps, update, nmon, prec)
// from the current step counter will replay: the same segment schedule,
// walked without running a step.
int Engine::prepare_graphs(int nsteps, int update, int nmon, int prec) {
    int r;
    if (nsteps < 0 || update <= 0) return JP_EINVAL;
    if (!use_graphs) return 0;
    if ((r = ensure_list(prec))) return r;
    if ((r = set_factors(-1))) return r;
    long long id = idloop;
    int step = 1;
    while (step <= nsteps) {
        const bool mon = nmon > 0 && (id + 1) % nmon == 0;
        int nplain = 0;
        if (!mon && !heating(id + 1)) {
            nplain = std::min(nsteps - step + 1, (int)(update - id % update));
            if (nmon > 0) nplain = std::min(nplain, (int)(nmon - (id + 1) % nmon));
        }
        if (nplain >= 2) {
            if (!step_graph(nplain, prec)) return 0;   // capture unavailable: the run launches kernels itself
            step += nplain;
            id += nplain;
        } else {
            ++step;
            ++id;
        }
    }
    return 0;
}


// ------------------------------------------------------------------ CMM

// CAL15M= CMM for the following evaluations and runs (one system): the cell
// levels from CELSIZ and CELLVL, raised to cover the extent of the current
// coordinates (incel.f90:29-57, 137-150), the far field every CMMUPD steps.
int Engine::set_cmm(int on, double cminsiz, int nlev, int cmmupd) {
    drop_graphs();
    have_list = false;
    if (!on) {
        if (cmm) {
            cmm = false;
            for (auto &h : hs)
                for (int k = 0; k < 3; ++k) h.dself[k] = h.dself_zd[k];
            if (alsd) {
                std::vector<double> ds(3 * (size_t)nsys);
                for (int s = 0; s < nsys; ++s)
                    for (int k = 0; k < 3; ++k) ds[3 * s + k] = hs[s].dself[k];
                CK(hipMemcpyAsync(d_dself.p, ds.data(), 8 * ds.size(), hipMemcpyHostToDevice, st));
                CK(hipStreamSynchronize(st));
            }
        }
        return 0;
    }
    if (nsys != 1 || !(cminsiz > 0.0) || !std::isfinite(cminsiz) || nlev < 1 || nlev > JP_CMM_MAXLEV || cmmupd < 1)
        return JP_EINVAL;
    std::vector<double> x(3 * (size_t)natom);
    CK(hipMemcpyAsync(x.data(), d_x.p, 8 * x.size(), hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    double ext = 0.0;
    for (int k = 0; k < 3; ++k) {
        double lo = x[k], hi = x[k];
        for (int i = 1; i < natom; ++i) {
            lo = std::min(lo, x[3 * (size_t)i + k]);
            hi = std::max(hi, x[3 * (size_t)i + k]);
        }
        ext = k == 0 ? hi - lo : std::max(hi - lo, ext);
    }
    if ((double)(1 << (nlev + 1)) * cminsiz < ext) {
        nlev = (int)std::ceil(std::log(ext / cminsiz) / std::log(2.0) - 1.0);
        if (nlev > JP_CMM_MAXLEV) return JP_EINVAL;   // md_run stops (incel.f90:46-52)
    }
    const HostSys &h = hs[0];
    CmmArgs A{};
    A.n = natom; A.nf = h.nfrag; A.nfrg = h.nfrg; A.nlev = nlev; A.ntyp = h.ntyp;
    for (int i = 0; i < nlev; ++i) {
        A.nv[nlev - 1 - i] = 4 << i;
        A.siz[i] = cminsiz * (double)(1 << i);
    }
    A.nv1 = A.nv[0];
    A.nc1 = A.nv1 * A.nv1 * A.nv1;
    A.rlim2 = h.cutoff * h.cutoff;
    A.idelay = (int)std::ceil(h.cutoff / A.siz[0]);
    if (A.idelay > 15) return JP_EINVAL;   // cells below cutoff / 15: beyond k_cmm_l2's run table
    A.ccoef = h.ccoef;
    const double vol = std::pow(4.0 * A.siz[0], 3.0);
    A.ipmax = (int)std::ceil(A.siz[0] * A.siz[0] * A.siz[0] * 0.27) * 2;
    A.capvd = (int)std::ceil(0.15 * vol) * 2;
    A.capel = A.capvd;
    A.capvw = (int)std::ceil(0.027 * vol) * 2;
    A.nkey = A.nc1 * A.nf * 2;
    std::vector<double> cg;
    int npc = 0;
    for (int l = 0; l < nlev; ++l) {
        A.cgoff[l] = (int)cg.size();
        A.pcoff[l] = npc;
        const double r = -0.5 * (A.nv[l] + 1);   // mk_c: centres siz (i - (nv + 1)/2), 1-based i
        for (int i = 1; i <= A.nv[l]; ++i) cg.push_back(A.siz[l] * ((double)i + r));
        npc += A.nv[l] * A.nv[l] * A.nv[l] * A.nf;
    }
    const size_t n = (size_t)natom;
    AL(d_cmm_cg, cg.size()); AL(d_cmm_c0, 3); AL(d_cmm_q, n); AL(d_cmm_sx0, 3 * n); AL(d_cmm_pot, 16 * (size_t)npc);
    AL(d_cmm_gnb, 9 * n); AL(d_cmm_enb, 6 * n); AL(d_cmm_gwk, 9 * n); AL(d_cmm_efar, 3 * n); AL(d_cmm_eatom, 6 * n);
    AL(d_cmm_gwkp, (size_t)CMM_FP * 12 * n);
    AL(d_cmm_wb, 12 * n); AL(d_cmm_wb2, 12 * n); AL(d_cmm_sx, n);
    AL(d_cmm_key, n); AL(d_cmm_kcnt, A.nkey); AL(d_cmm_kstart, A.nkey + 1); AL(d_cmm_kfill, A.nkey);
    AL(d_cmm_tmp, n); AL(d_cmm_sorig, n); AL(d_cmm_spos, n); AL(d_cmm_sflag, n); AL(d_cmm_smol, n);
    AL(d_cmm_npc, npc); AL(d_cmm_err, 2);
    AL(d_cmm_xmask, n); AL(d_cmm_xwide, n); AL(d_cmm_l2cnt, n); AL(d_cmm_l2ptr, n + 1); AL(d_cmm_l2idx, 1);
    AL(d_cmm_vflag, n); AL(d_cmm_vpos, n + 1); AL(d_cmm_vlist, n);
    AL(d_cmm_kvcnt, A.nkey); AL(d_cmm_kvstart, A.nkey + 1); AL(d_cmm_vx0, n); AL(d_cmm_vf, n); AL(d_cmm_va, n);
    AL(d_cmm_vm, n); AL(d_cmm_vp, n);
    AL(d_cmm_tsum, 2 * ((std::max((size_t)A.nkey, n) + 1023) / 1024 + 1));
    CK(hipMemcpyAsync(d_cmm_cg.p, cg.data(), 8 * cg.size(), hipMemcpyHostToDevice, st));
    CK(hipMemcpyAsync(d_cmm_q.p, charge_h.data(), 8 * n, hipMemcpyHostToDevice, st));
    CK(hipMemsetAsync(d_cmm_gwk.p, 0, 72 * n, st));
    CK(hipMemsetAsync(d_cmm_efar.p, 0, 24 * n, st));
    // slots are atoms, with no periodic images
    std::vector<int> id(n);
    std::iota(id.begin(), id.end(), 0);
    CK(hipMemcpyAsync(d_inum.p, id.data(), 4 * n, hipMemcpyHostToDevice, st));
    CK(hipMemsetAsync(d_img.p, 0, 12 * n, st));
    A.cg = d_cmm_cg.p; A.c0 = d_cmm_c0.p; A.x = d_x.p; A.chg = d_chg.p; A.q = d_cmm_q.p; A.tflag = d_tflag.p;
    A.mol = d_mol.p; A.xptr = d_xptr.p; A.xidx = d_xidx.p; A.lj = d_lj64.p;
    A.key = d_cmm_key.p; A.kcnt = d_cmm_kcnt.p; A.kstart = d_cmm_kstart.p; A.kfill = d_cmm_kfill.p;
    A.tmp = d_cmm_tmp.p; A.sorig = d_cmm_sorig.p; A.spos = d_cmm_spos.p; A.sflag = d_cmm_sflag.p;
    A.smol = d_cmm_smol.p; A.sx = d_cmm_sx.p; A.sx0 = d_cmm_sx0.p; A.npc = d_cmm_npc.p; A.pot = d_cmm_pot.p;
    A.err = d_cmm_err.p; A.sp = d_sp.p; A.efac = d_efac.p; A.gnb = d_cmm_gnb.p; A.enb = d_cmm_enb.p;
    A.gwk = d_cmm_gwk.p; A.efar = d_cmm_efar.p; A.gslot = d_gslot.p; A.eatom = d_cmm_eatom.p;
    A.gwkp = d_cmm_gwkp.p;
    A.xmask = d_cmm_xmask.p; A.xwide = d_cmm_xwide.p; A.l2cnt = d_cmm_l2cnt.p; A.l2ptr = d_cmm_l2ptr.p;
    A.l2idx = d_cmm_l2idx.p; A.vflag = d_cmm_vflag.p; A.vpos = d_cmm_vpos.p; A.vlist = d_cmm_vlist.p;
    A.kvcnt = d_cmm_kvcnt.p; A.kvstart = d_cmm_kvstart.p; A.vx0 = d_cmm_vx0.p; A.vf = d_cmm_vf.p;
    A.va = d_cmm_va.p; A.vm = d_cmm_vm.p; A.vp = d_cmm_vp.p;
    cm = A;
    hipLaunchKernelGGL(k_cmm_xmask, grid1(n, 256), 256, 0, st, A);
    CK(hipGetLastError());
    cmm_vcap = 0;
    hs[0].dself[0] = hs[0].dself[1] = hs[0].dself[2] = 0.0;   // no ZD self energy under CMM
    if (alsd) {
        const double z[3] = {0.0, 0.0, 0.0};
        CK(hipMemcpyAsync(d_dself.p, z, 24, hipMemcpyHostToDevice, st));
    }
    CK(hipStreamSynchronize(st));
    cmm = true;
    cmm_upd = cmmupd;
    cmm_itwin = true;
    list_half = false;
    return 0;
}

// celprp under CMM: cell centres from the coordinate mean, binning and the
// sort, cell counts and multipoles of every level, the table checks; one
// read-back of the status (ier 28, 51, 99).
int Engine::cmm_celprp() {
    CmmArgs &A = cm;
    const unsigned gn = grid1(natom, 256);
    hipLaunchKernelGGL(k_cmm_mean, 1, 1024, 0, st, A);
    hipLaunchKernelGGL(k_cmm_reset, grid1(std::max(A.nkey, 2), 256), 256, 0, st, A);
    hipLaunchKernelGGL(k_cmm_bin, gn, 256, 0, st, A);
    cmm_scan(A.nkey, A.kcnt, A.kstart);
    hipLaunchKernelGGL(k_cmm_place, gn, 256, 0, st, A);
    hipLaunchKernelGGL(k_cmm_rank, gn, 256, 0, st, A);
    hipLaunchKernelGGL(k_cmm_fill, gn, 256, 0, st, A);
    hipLaunchKernelGGL(k_cmm_kv, grid1(A.nkey, 256), 256, 0, st, A);
    cmm_scan(A.nkey, A.kvcnt, A.kvstart);
    hipLaunchKernelGGL(k_cmm_vcopy, gn, 256, 0, st, A);
    for (int l = 0; l < A.nlev; ++l) {
        const int nq = A.nv[l] * A.nv[l] * A.nv[l] * A.nf;
        hipLaunchKernelGGL(k_cmm_npc, grid1(nq, 256), 256, 0, st, A, l);
    }
    for (int l = 0; l < A.nlev; ++l) {
        const int nq = A.nv[l] * A.nv[l] * A.nv[l] * A.nf;
        hipLaunchKernelGGL(k_cmm_pol, grid1(nq, 128), 128, 0, st, A, l);
    }
    hipLaunchKernelGGL(k_cmm_check, gn, 256, 0, st, A);
    if (A.idelay > 1) hipLaunchKernelGGL(k_cmm_l2<false>, A.nc1, 256, 0, st, A);
    // the second-nearest LJ tables: lengths -> offsets, the atoms that have one
    cmm_scan(natom, A.l2cnt, A.l2ptr);
    hipLaunchKernelGGL(k_cmm_vflag, gn, 256, 0, st, A);
    cmm_scan(natom, A.vflag, A.vpos);
    CK(hipGetLastError());
    int st4[4] = {0, 0, 0, 0};
    CK(hipMemcpyAsync(st4, d_cmm_err.p, 8, hipMemcpyDeviceToHost, st));
    CK(hipMemcpyAsync(st4 + 2, A.l2ptr + natom, 4, hipMemcpyDeviceToHost, st));
    CK(hipMemcpyAsync(st4 + 3, A.vpos + natom, 4, hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    if (st4[0] & 1) return JP_IER_CELL;    // celpar.f90:172-192
    if (st4[0] & 2) return JP_IER_LIST;    // neartb.f90:368-400
    if (st4[1] != 2 * hs[0].nexcl) return JP_IER_EXCL;   // neartb.f90:446-453
    if ((size_t)st4[2] > d_cmm_l2idx.n) {   // grows only: captured step graphs hold the pointer
        drop_graphs();
        AL(d_cmm_l2idx, (size_t)st4[2] + (size_t)st4[2] / 4);
        cm.l2idx = d_cmm_l2idx.p;
    }
    if (st4[3] > cmm_vcap) {
        drop_graphs();
        cmm_vcap = st4[3] + st4[3] / 8;
    }
    if (A.idelay > 1) hipLaunchKernelGGL(k_cmm_l2<true>, A.nc1, 256, 0, st, cm);
    hipLaunchKernelGGL(k_cmm_vfill, gn, 256, 0, st, cm);
    CK(hipGetLastError());
    return 0;
}

// k_cmm_far's grid: CMM_FP blocks per group of 2 x 2 x 2 level-1 cells (per
// level-1 cell with one level).
static unsigned cmm_far_blocks(const CmmArgs &A) {
    const int ng = A.nlev > 1 ? A.nv1 / 2 : A.nv1;
    return (unsigned)(ng * ng * ng * CMM_FP);
}

// Exclusive scan of n ints (out[n] the total) in three launches.
int Engine::cmm_scan(int n, const int *in, int *out) {
    const int nt = (n + 1023) / 1024;
    int *t0 = d_cmm_tsum.p, *t1 = d_cmm_tsum.p + (d_cmm_tsum.n / 2);
    hipLaunchKernelGGL(k_scan_tile, std::max(nt, 1), 256, 0, st, n, in, out, t0);
    hipLaunchKernelGGL(k_scan, 1, 1024, 0, st, nt, (const int *)t0, t1);
    hipLaunchKernelGGL(k_scan_add, grid1((size_t)n + 1, 256), 256, 0, st, n, out, (const int *)t1, nt);
    CK(hipGetLastError());
    return 0;
}

template <bool MULTI, bool E>
static void launch_cmm(Engine &g, bool itwin) {
    const CmmArgs &A = g.cm;
    hipLaunchKernelGGL(k_cmm_gather, grid1(A.n, 256), 256, 0, g.st, A);
    hipLaunchKernelGGL((k_cmm_near<MULTI, E>), A.nc1, CMM_NT, sizeof(double2) * A.ntyp * A.ntyp, g.st, A);
    if (A.idelay > 1 && g.cmm_vcap > 0)
        hipLaunchKernelGGL((k_cmm_lj2<MULTI, E>), grid1(g.cmm_vcap, 4), 256, 0, g.st, A);
    if (itwin) {
        hipLaunchKernelGGL(k_cmm_far<MULTI>, cmm_far_blocks(A), CMM_NT, 0, g.st, A);
        hipLaunchKernelGGL(k_cmm_farsum, grid1(A.n, 256), 256, 0, g.st, A);
    }
    hipLaunchKernelGGL(k_cmm_combine<E>, grid1(A.n, 256), 256, 0, g.st, A);
}

// enear_CMM_doub and the far field (when itwin) into d_gslot; with energy the
// per-atom energies of the step.
int Engine::cmm_forces(bool energy, bool itwin) {
    const bool multi = hs[0].nfrg > 1;
    if (multi) {
        if (energy) launch_cmm<true, true>(*this, itwin);
        else launch_cmm<true, false>(*this, itwin);
    } else {
        if (energy) launch_cmm<false, true>(*this, itwin);
        else launch_cmm<false, false>(*this, itwin);
    }
    CK(hipGetLastError());
    return 0;
}

// Mean durations of the near-field kernel (k_cmm_near, with k_cmm_lj2 when
// idelay > 1) and of the far-field kernel over reps launches each, between
// HIP events on the engine's stream, at the device coordinates.
int Engine::cmm_time(int reps, double &near, double &far) {
    int r;
    if ((r = ensure_list(JP_PREC_FP64))) return r;
    if ((r = set_factors(-1))) return r;
    if ((r = cmm_forces(false, true))) return r;
    const CmmArgs &A = cm;
    const bool multi = hs[0].nfrg > 1;
    CK(hipEventRecord(ev[0], st));
    const size_t lds = sizeof(double2) * A.ntyp * A.ntyp;
    for (int k = 0; k < reps; ++k) {
        if (multi) hipLaunchKernelGGL((k_cmm_near<true, false>), A.nc1, CMM_NT, lds, st, A);
        else hipLaunchKernelGGL((k_cmm_near<false, false>), A.nc1, CMM_NT, lds, st, A);
        if (A.idelay > 1 && cmm_vcap > 0) {
            if (multi) hipLaunchKernelGGL((k_cmm_lj2<true, false>), grid1(cmm_vcap, 4), 256, 0, st, A);
            else hipLaunchKernelGGL((k_cmm_lj2<false, false>), grid1(cmm_vcap, 4), 256, 0, st, A);
        }
    }
    CK(hipEventRecord(ev[1], st));
    for (int k = 0; k < reps; ++k) {
        if (multi) hipLaunchKernelGGL(k_cmm_far<true>, cmm_far_blocks(A), CMM_NT, 0, st, A);
        else hipLaunchKernelGGL(k_cmm_far<false>, cmm_far_blocks(A), CMM_NT, 0, st, A);
        hipLaunchKernelGGL(k_cmm_farsum, grid1(A.n, 256), 256, 0, st, A);
    }
    CK(hipEventRecord(ev[2], st));
    CK(hipGetLastError());
    CK(hipEventSynchronize(ev[2]));
    float m1 = 0, m2 = 0;
    CK(hipEventElapsedTime(&m1, ev[0], ev[1]));
    CK(hipEventElapsedTime(&m2, ev[1], ev[2]));
    near = m1 * 1e-3 / reps;
    far = m2 * 1e-3 / reps;
    // the timing launches recomputed the same near and far sums: the state is unchanged
    return cmm_forces(false, true);
}

__global__ void k_cmm_zero(double *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// The step's per-atom CMM energies summed per system in a fixed order into
// the pair-energy slots of d_rsum (halved with them); no excess energy.
int Engine::cmm_sums() {
    hipLaunchKernelGGL(k_cmm_zero, 1, 64, 0, st, d_rsum.p + 6 * (size_t)nsys, 3 * nsys);
    CK(hipGetLastError());
    return seg_sum<double>(d_abeg.p, d_aend.p, 6, d_cmm_eatom.p, d_part.p, d_rsum.p);
}

int Engine::get_list(int32_t *cell3, int64_t *npairs, int32_t *ptr, int32_t *idx, int64_t cap) {
    if (!have_list) return JP_EINVAL;
    if (cmm) {   // CMM keeps no pair list: the level-1 cell (1-based per axis) of every atom
        if (ptr || idx) return JP_EINVAL;
        std::vector<int> key(natom);
        CK(hipMemcpyAsync(key.data(), d_cmm_key.p, 4 * (size_t)natom, hipMemcpyDeviceToHost, st));
        CK(hipStreamSynchronize(st));
        const int nv = cm.nv1;
        for (int a = 0; a < natom && cell3; ++a) {
            const int c = key[a] / (2 * cm.nf);
            cell3[3 * a] = c % nv + 1;
            cell3[3 * a + 1] = (c / nv) % nv + 1;
            cell3[3 * a + 2] = c / (nv * nv) + 1;
        }
        if (npairs) *npairs = 0;
        return 0;
    }
    CK(hipStreamSynchronize(st));
    std::vector<int> perm(nslot);
    std::vector<int2> lc(nclu);
    std::vector<uint2> ent((size_t)nclu * ccap);
    CK(hipMemcpy(lc.data(), d_lcnt.p, 8 * lc.size(), hipMemcpyDeviceToHost));
    CK(hipMemcpy(perm.data(), d_perm.p, 4 * (size_t)nslot, hipMemcpyDeviceToHost));
    CK(hipMemcpy(ent.data(), d_ent.p, sizeof(uint2) * ent.size(), hipMemcpyDeviceToHost));
    if (cell3) CK(hipMemcpy(cell3, d_refcell.p, 12 * (size_t)natom, hipMemcpyDeviceToHost));
    std::vector<std::vector<int>> nb(natom);
    for (int c = 0; c < nclu; ++c) {
        const size_t e0 = (size_t)c * ccap;
        const int cp = ccap;
        for (int e = 0; e < lc[c].x + lc[c].y; ++e) {
            const uint2 E = ent[e < lc[c].x ? e0 + e : e0 + cp - lc[c].y + (e - lc[c].x)];
            const int aj = perm[E.x & JP_SLOT_MASK];
            for (int k = 0; k < JP_CL; ++k) {
                if (!((E.y >> k) & 1u)) continue;
                const int ai = perm[c * JP_CL + k];
                if (ai >= 0 && aj >= 0) {
                    nb[ai].push_back(aj);
                    if (list_half) nb[aj].push_back(ai);   // each unique pair once: both directions
                }
            }
        }
    }
    int64_t tot = 0;
    for (auto &v : nb) tot += (int64_t)v.size();
    *npairs = tot;
    if (!idx) return 0;
    if (tot > cap) return JP_EINVAL;
    int64_t w = 0;
    for (int i = 0; i < natom; ++i) {
        ptr[i] = (int32_t)w;
        std::sort(nb[i].begin(), nb[i].end());
        int base = sp[atom_sys_h[i]].atom0;
        for (int j : nb[i]) idx[w++] = j - base + 1;
    }
    ptr[natom] = (int32_t)w;
    return 0;
}

// jp_count_pairs without allocation: per-cluster counts into the static
// scratch, summed by one block in cluster order.
int Engine::count_pairs_noalloc(int64_t *np, int64_t *nlj, int64_t *nlist) {
    if (!have_list || cmm) return JP_EINVAL;
    if (nclu > JP_NOALLOC_MAX_CLU) return JP_EINVAL;   // beyond the scratch
    unsigned long long *sc = noalloc_scratch();
    if (!sc) return JP_ENODEV;
    unsigned long long *tot = sc + 3 * (size_t)JP_NOALLOC_MAX_CLU;
    hipLaunchKernelGGL(k_count_pairs_part, nclu, 64, 0, st, nclu, d_clu_sys.p, d_sp.p, d_ys.p, ccap, d_lcnt.p,
                       d_ent.p, sc);
    hipLaunchKernelGGL(k_sum3, 1, 256, 0, st, nclu, (const unsigned long long *)sc, tot);
    CK(hipGetLastError());
    unsigned long long c[3] = {0, 0, 0};
    CK(hipMemcpyAsync(c, tot, 24, hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    if (np) *np = (int64_t)c[0];
    if (nlj) *nlj = (int64_t)c[1];
    if (nlist) *nlist = (int64_t)c[2];
    return 0;
}

int Engine::count_pairs(int64_t *np, int64_t *nlj, int64_t *nlist) {
    if (!have_list || cmm) return JP_EINVAL;
    CK(hipMemsetAsync(d_u64.p, 0, 24, st));
    hipLaunchKernelGGL(k_count_pairs, nclu, 64, 0, st, nclu, d_clu_sys.p, d_sp.p, d_ys.p, ccap, d_lcnt.p,
                       d_ent.p, d_u64.p);
    unsigned long long c[3] = {0, 0, 0};
    CK(hipMemcpyAsync(c, d_u64.p, 24, hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    if (np) *np = (int64_t)c[0];
    if (nlj) *nlj = (int64_t)c[1];
    if (nlist) *nlist = (int64_t)c[2];
    return 0;
}

}  // namespace

struct jp_ctx {
    Engine e;
};
struct jp_batch {
    Engine e;
};

extern "C" {

const char *jp_version(void) { return "jpgpu 0.1 (gfx950)"; }

int jp_ctx_create(const jp_sys_desc *desc, jp_ctx **out) {
    if (!desc || !out) return JP_EINVAL;
    jp_ctx *c = new jp_ctx();
    int r = c->e.init(desc, 1);
    if (r) {
        delete c;
        *out = nullptr;
        return r;
    }
    *out = c;
    return 0;
}

int jp_ctx_destroy(jp_ctx *ctx) {
    delete ctx;
    return 0;
}

int jp_set_state(jp_ctx *ctx, const double *cord, const double *vel) {
    if (!ctx) return JP_EINVAL;
    Engine &E = ctx->e;
    if (cord) E.cmm_itwin = true;   // CMM: a new start computes the far field first (incel's itwin = 1)
    if (cord) CK(hipMemcpyAsync(E.d_x.p, cord, 24 * (size_t)E.natom, hipMemcpyHostToDevice, E.st));
    if (vel) CK(hipMemcpyAsync(E.d_v.p, vel, 24 * (size_t)E.natom, hipMemcpyHostToDevice, E.st));
    if (cord && E.have_list) {
        int r = E.refresh_sorted();
        if (r) return r;
    }
    CK(hipStreamSynchronize(E.st));
    return 0;
}

int jp_get_state(jp_ctx *ctx, double *cord, double *vel) {
    if (!ctx) return JP_EINVAL;
    Engine &E = ctx->e;
    if (cord) CK(hipMemcpyAsync(cord, E.d_x.p, 24 * (size_t)E.natom, hipMemcpyDeviceToHost, E.st));
    if (vel) CK(hipMemcpyAsync(vel, E.d_v.p, 24 * (size_t)E.natom, hipMemcpyDeviceToHost, E.st));
    CK(hipStreamSynchronize(E.st));
    return 0;
}

int jp_build_list(jp_ctx *ctx, const double *cord) {
    if (!ctx) return JP_EINVAL;
    Engine &E = ctx->e;
    if (cord) CK(hipMemcpyAsync(E.d_x.p, cord, 24 * (size_t)E.natom, hipMemcpyHostToDevice, E.st));
    return E.build_list();
}

int jp_get_list(jp_ctx *ctx, int32_t *cell_of_atom, int64_t *npairs, int32_t *ptr, int32_t *idx, int64_t cap) {
    if (!ctx || !npairs) return JP_EINVAL;
    return ctx->e.get_list(cell_of_atom, npairs, ptr, idx, cap);
}

int jp_nonbond(jp_ctx *ctx, const double *cord, double *grad, double *eelec, double *evdw, double *ehyd, int precision) {
    if (!ctx || (precision != JP_PREC_FP64 && precision != JP_PREC_FP32_FIXED)) return JP_EINVAL;
    int r = ctx->e.nonbond(cord, grad, eelec, evdw, precision);
    if (ehyd)
        for (int k = 0; k < ctx->e.hs[0].nfrg; ++k) ehyd[k] = 0.0;   // no 12-10 H-bond pairs (iyppid != 2)
    return r;
}

int jp_bonded(jp_ctx *ctx, const double *cord, double *grad, double *eterm) {
    if (!ctx) return JP_EINVAL;
    return ctx->e.bonded(cord, grad, eterm);
}

int jp_shake(jp_ctx *ctx, double *cord, const double *cord_old) {
    if (!ctx || !cord || !cord_old) return JP_EINVAL;
    Engine &E = ctx->e;
    CK(hipMemcpyAsync(E.d_x.p, cord, 24 * (size_t)E.natom, hipMemcpyHostToDevice, E.st));
    CK(hipMemcpyAsync(E.d_xold.p, cord_old, 24 * (size_t)E.natom, hipMemcpyHostToDevice, E.st));
    CK(hipMemsetAsync(E.d_err.p, 0, 4, E.st));
    int ier = 0, r = E.shake(ier);
    if (r) return r;
    int err = 0;
    CK(hipMemcpyAsync(cord, E.d_x.p, 24 * (size_t)E.natom, hipMemcpyDeviceToHost, E.st));
    CK(hipMemcpyAsync(&err, E.d_err.p, 4, hipMemcpyDeviceToHost, E.st));
    CK(hipStreamSynchronize(E.st));
    return step_ier(err);
}

int jp_run_nve(jp_ctx *ctx, int32_t nsteps, int32_t update, int32_t nmon, int32_t precision, jp_energy_log *log) {
    if (!ctx || (precision != JP_PREC_FP64 && precision != JP_PREC_FP32_FIXED)) return JP_EINVAL;
    if (nmon > 0 && !log) return JP_EINVAL;
    return ctx->e.run_nve(nsteps, update, nmon, precision, log);
}

int jp_restart_step(jp_ctx *ctx, int32_t precision, jp_energy_log *log) {
    if (!ctx || (precision != JP_PREC_FP64 && precision != JP_PREC_FP32_FIXED)) return JP_EINVAL;
    return ctx->e.restart_step(precision, log);
}

int jp_batch_restart_step(jp_batch *b, int32_t precision, jp_energy_log *log) {
    if (!b || (precision != JP_PREC_FP64 && precision != JP_PREC_FP32_FIXED)) return JP_EINVAL;
    return b->e.restart_step(precision, log);
}

int jp_last_timing(jp_ctx *ctx, double *t_total, double *t_pair, double *t_list, double *t_step) {
    if (!ctx) return JP_EINVAL;
    if (t_total) *t_total = ctx->e.t_total;
    // mean pair-kernel time over the host-launched steps, times all steps
    if (t_pair) *t_pair = ctx->e.n_pair ? ctx->e.t_pair / ctx->e.n_pair * ctx->e.n_pair_steps : 0.0;
    if (t_list) *t_list = ctx->e.t_list;
    if (t_step) *t_step = ctx->e.t_step;
    return 0;
}

int jp_time_pair(jp_ctx *ctx, int32_t precision, int32_t reps, double *seconds) {
    if (!ctx || !seconds || reps <= 0 || (precision != JP_PREC_FP64 && precision != JP_PREC_FP32_FIXED))
        return JP_EINVAL;
    return ctx->e.time_pair(precision, reps, *seconds);
}

int jp_prepare_graphs(jp_ctx *ctx, int32_t nsteps, int32_t update, int32_t nmon, int32_t precision) {
    if (!ctx || (precision != JP_PREC_FP64 && precision != JP_PREC_FP32_FIXED)) return JP_EINVAL;
    return ctx->e.prepare_graphs(nsteps, update, nmon, precision);
}

int jp_batch_prepare_graphs(jp_batch *b, int32_t nsteps, int32_t update, int32_t nmon, int32_t precision) {
    if (!b || (precision != JP_PREC_FP64 && precision != JP_PREC_FP32_FIXED)) return JP_EINVAL;
    return b->e.prepare_graphs(nsteps, update, nmon, precision);
}

int jp_set_step(jp_ctx *ctx, int64_t idloop) {
    if (!ctx || idloop < 0) return JP_EINVAL;
    ctx->e.idloop = idloop;
    return 0;
}

int jp_get_step(jp_ctx *ctx, int64_t *idloop) {
    if (!ctx || !idloop) return JP_EINVAL;
    *idloop = ctx->e.idloop;
    return 0;
}

int jp_batch_set_step(jp_batch *b, int64_t idloop) {
    if (!b || idloop < 0) return JP_EINVAL;
    b->e.idloop = idloop;
    return 0;
}

int jp_count_pairs_noalloc(jp_ctx *ctx, int64_t *npairs, int64_t *nlj, int64_t *nlisted) {
    if (!ctx || !npairs) return JP_EINVAL;
    return ctx->e.count_pairs_noalloc(npairs, nlj, nlisted);
}

int jp_count_pairs(jp_ctx *ctx, int64_t *npairs, int64_t *nlj, int64_t *nlisted) {
    if (!ctx || !npairs) return JP_EINVAL;
    return ctx->e.count_pairs(npairs, nlj, nlisted);
}

int jp_batch_create(const jp_sys_desc *descs, int32_t nsys, jp_batch **out) {
    if (!descs || !out || nsys <= 0) return JP_EINVAL;
    jp_batch *b = new jp_batch();
    int r = b->e.init(descs, nsys);
    if (r) {
        delete b;
        *out = nullptr;
        return r;
    }
    *out = b;
    return 0;
}

int jp_batch_run_nve(jp_batch *b, int32_t nsteps, int32_t update, int32_t nmon, int32_t precision, jp_energy_log *log) {
    if (!b || (precision != JP_PREC_FP64 && precision != JP_PREC_FP32_FIXED)) return JP_EINVAL;
    if (nmon > 0 && !log) return JP_EINVAL;
    return b->e.run_nve(nsteps, update, nmon, precision, log);
}

int jp_batch_get_state(jp_batch *b, int32_t isys, double *cord, double *vel) {
    if (!b || isys < 0 || isys >= b->e.nsys) return JP_EINVAL;
    Engine &E = b->e;
    size_t off = 3 * (size_t)E.sp[isys].atom0, n = 3 * (size_t)E.sp[isys].natom;
    if (cord) CK(hipMemcpyAsync(cord, E.d_x.p + off, 8 * n, hipMemcpyDeviceToHost, E.st));
    if (vel) CK(hipMemcpyAsync(vel, E.d_v.p + off, 8 * n, hipMemcpyDeviceToHost, E.st));
    CK(hipStreamSynchronize(E.st));
    return 0;
}

int jp_batch_set_state(jp_batch *b, int32_t isys, const double *cord, const double *vel) {
    if (!b || isys < 0 || isys >= b->e.nsys) return JP_EINVAL;
    Engine &E = b->e;
    size_t off = 3 * (size_t)E.sp[isys].atom0, n = 3 * (size_t)E.sp[isys].natom;
    if (cord) CK(hipMemcpyAsync(E.d_x.p + off, cord, 8 * n, hipMemcpyHostToDevice, E.st));
    if (vel) CK(hipMemcpyAsync(E.d_v.p + off, vel, 8 * n, hipMemcpyHostToDevice, E.st));
    if (cord && E.have_list) {
        int r = E.refresh_sorted();
        if (r) return r;
    }
    CK(hipStreamSynchronize(E.st));
    return 0;
}

int jp_batch_last_timing(jp_batch *b, double *t_total) {
    if (!b) return JP_EINVAL;
    if (t_total) *t_total = b->e.t_total;
    return 0;
}

static int set_option(Engine &E, int32_t option, int32_t value) {
    switch (option) {
        case JP_OPT_GRAPHS:
            E.use_graphs = value ? 1 : 0;
            if (!value) E.drop_graphs();
            return 0;
        case JP_OPT_AUTOTUNE:
            E.autotune = value ? 1 : 0;
            E.tuned = 0;
            return 0;
        case JP_OPT_PAIR_WAVES:
            if (value != 1 && value != 2 && value != 4 && value != 8) return JP_EINVAL;
            E.pw = std::max(2, value);   // full-list kernel: 2, 4, 8
            E.pwh = value;               // half-list kernel: 1, 2, 4, 8
            E.autotune = 0;
            return 0;
        case JP_OPT_HALF_LIST:
            E.half_opt = value ? 1 : 0;
            E.have_list = false;
            E.list_half = E.want_half(JP_PREC_FP32_FIXED);
            return 0;
        case JP_OPT_LIST_WAVES:
            if (value != 1 && value != 2 && value != 4 && value != 8) return JP_EINVAL;
            E.lw = value;
            E.autotune = 0;
            return 0;
        case JP_OPT_LIST_SEGMENT:
            if (value < 64 || value % 64) return JP_EINVAL;   // whole 64-entry chunks
            E.ccap_force = value;
            return 0;
        case JP_OPT_CELL_TABLE:
            E.cell_table = value ? 1 : 0;
            E.have_list = false;   // the next run or call builds the list again
            E.list_half = E.want_half(JP_PREC_FP32_FIXED);
            return 0;
        case JP_OPT_GRID_SHIFT:
            if (value < -3 || value > 3) return JP_EINVAL;
            E.autotune = 0;
            return value == E.grid_shift ? 0 : E.regrid(value);
        case JP_OPT_LIST_BATCH:
            if (value != 1 && value != 2 && value != 4) return JP_EINVAL;
            E.lnb = value;
            E.autotune = 0;
            return 0;
        default:
            return JP_EINVAL;
    }
}

static int get_option(const Engine &E, int32_t option, int32_t *value) {
    if (!value) return JP_EINVAL;
    switch (option) {
        case JP_OPT_GRAPHS: *value = E.use_graphs; return 0;
        case JP_OPT_AUTOTUNE: *value = E.autotune; return 0;
        case JP_OPT_PAIR_WAVES: *value = E.want_half(JP_PREC_FP32_FIXED) ? E.pwh : E.pw; return 0;
        case JP_OPT_HALF_LIST: *value = E.half_opt; return 0;
        case JP_OPT_LIST_WAVES: *value = E.lw; return 0;
        case JP_OPT_CELL_TABLE: *value = E.cell_table; return 0;
        case JP_OPT_GRAPH_CAPTURES: *value = E.captures_last_run; return 0;
        case JP_OPT_GRID_SHIFT: *value = E.grid_shift; return 0;
        case JP_OPT_LIST_BATCH: *value = E.lnb; return 0;
        default: return JP_EINVAL;
    }
}

int jp_get_option(jp_ctx *ctx, int32_t option, int32_t *value) {
    if (!ctx) return JP_EINVAL;
    return get_option(ctx->e, option, value);
}

int jp_batch_get_option(jp_batch *b, int32_t option, int32_t *value) {
    if (!b) return JP_EINVAL;
    return get_option(b->e, option, value);
}

int jp_set_option(jp_ctx *ctx, int32_t option, int32_t value) {
    if (!ctx) return JP_EINVAL;
    return set_option(ctx->e, option, value);
}

int jp_batch_set_option(jp_batch *b, int32_t option, int32_t value) {
    if (!b) return JP_EINVAL;
    return set_option(b->e, option, value);
}

int jp_set_mcmd(jp_ctx *ctx, int32_t on, int32_t nwindow, int32_t ndeg, const double *c, const double *low,
                const double *high, double alphalw, double alphaup, double celw, double ceup, double tempce) {
    if (!ctx) return JP_EINVAL;
    return ctx->e.set_mcmd(on, nwindow, ndeg, c, low, high, alphalw, alphaup, celw, ceup, tempce);
}

int jp_set_heating(jp_ctx *ctx, int32_t heatlo, double startt) {
    if (!ctx) return JP_EINVAL;
    return ctx->e.set_heating(heatlo, startt);
}

int jp_batch_set_heating(jp_batch *b, int32_t heatlo, double startt) {
    if (!b) return JP_EINVAL;
    return b->e.set_heating(heatlo, startt);
}

int jp_set_stopce(jp_ctx *ctx, int32_t mode, int32_t natom, int32_t nkind) {
    if (!ctx) return JP_EINVAL;
    return ctx->e.set_stopce(mode, &natom, &nkind);
}

int jp_batch_set_stopce(jp_batch *b, int32_t mode, const int32_t *natom, const int32_t *nkind) {
    if (!b) return JP_EINVAL;
    return b->e.set_stopce(mode, natom, nkind);
}

int jp_start_step(jp_ctx *ctx, int32_t precision) {
    if (!ctx || (precision != JP_PREC_FP64 && precision != JP_PREC_FP32_FIXED)) return JP_EINVAL;
    return ctx->e.start_step(precision);
}

int jp_batch_start_step(jp_batch *b, int32_t precision) {
    if (!b || (precision != JP_PREC_FP64 && precision != JP_PREC_FP32_FIXED)) return JP_EINVAL;
    return b->e.start_step(precision);
}

int jp_time_cmm(jp_ctx *ctx, int32_t reps, double *near, double *far) {
    if (!ctx || reps < 1 || !ctx->e.cmm) return JP_EINVAL;
    double a = 0, b = 0;
    const int r = ctx->e.cmm_time(reps, a, b);
    if (near) *near = a;
    if (far) *far = b;
    return r;
}

int jp_set_cmm(jp_ctx *ctx, int32_t on, double cminsiz, int32_t nlev, int32_t cmmupd) {
    if (!ctx) return JP_EINVAL;
    return ctx->e.set_cmm(on, cminsiz, nlev, cmmupd);
}

int jp_set_alsd(jp_ctx *ctx, int32_t on, int32_t nwindow, int32_t ndeg, const double *c, const double *low,
                const double *high, double lower, double upper, double celw, double ceup, double tempce,
                double lambda_m, int32_t temp_control) {
    if (!ctx) return JP_EINVAL;
    return ctx->e.set_alsd(on, nwindow, ndeg, c, low, high, lower, upper, celw, ceup, tempce, lambda_m, temp_control);
}

int jp_batch_set_alsd(jp_batch *b, int32_t on, int32_t nwindow, int32_t ndeg, const double *c, const double *low,
                      const double *high, double lower, double upper, double celw, double ceup, double tempce,
                      double lambda_m, int32_t temp_control) {
    if (!b) return JP_EINVAL;
    return b->e.set_alsd(on, nwindow, ndeg, c, low, high, lower, upper, celw, ceup, tempce, lambda_m, temp_control);
}

static int get_lambda(Engine &E, int s, double *lambda, double *lambda_v) {
    if (s < 0 || s >= E.nsys) return JP_EINVAL;
    if (lambda) *lambda = E.hs[s].lambda;
    if (lambda_v) *lambda_v = E.hs[s].lambda_v;
    return 0;
}

static int set_lambda(Engine &E, int s, double lambda, double lambda_v) {
    if (s < 0 || s >= E.nsys || !std::isfinite(lambda) || !std::isfinite(lambda_v)) return JP_EINVAL;
    if (E.alsd) {
        int r = E.lam_download();   // the other systems' state
        if (r) return r;
    }
    E.hs[s].lambda = lambda;
    E.hs[s].lambda_v = lambda_v;
    int r = E.set_factors(-1);
    if (!r && E.alsd) r = E.lam_upload();
    return r;
}

int jp_get_lambda(jp_ctx *ctx, double *lambda, double *lambda_v) {
    if (!ctx) return JP_EINVAL;
    return get_lambda(ctx->e, 0, lambda, lambda_v);
}

int jp_set_lambda(jp_ctx *ctx, double lambda, double lambda_v) {
    if (!ctx) return JP_EINVAL;
    return set_lambda(ctx->e, 0, lambda, lambda_v);
}

int jp_batch_get_lambda(jp_batch *b, int32_t isys, double *lambda, double *lambda_v) {
    if (!b) return JP_EINVAL;
    return get_lambda(b->e, isys, lambda, lambda_v);
}

int jp_batch_set_lambda(jp_batch *b, int32_t isys, double lambda, double lambda_v) {
    if (!b) return JP_EINVAL;
    return set_lambda(b->e, isys, lambda, lambda_v);
}

int jp_set_canonical(jp_ctx *ctx, int32_t on, double temps, int32_t icslop, double fcstol, const int32_t *group) {
    if (!ctx) return JP_EINVAL;
    return ctx->e.set_canonical(on, temps, icslop, fcstol, group);
}

int jp_batch_set_canonical(jp_batch *b, int32_t on, double temps, int32_t icslop, double fcstol,
                           const int32_t *group) {
    if (!b) return JP_EINVAL;
    return b->e.set_canonical(on, temps, icslop, fcstol, group);
}

int jp_batch_destroy(jp_batch *b) {
    delete b;
    return 0;
}

}  // extern "C"
