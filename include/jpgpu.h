/*
 * jpgpu.h - C ABI of the MI355X (gfx950) implementation of j_presto md_run's
 * zero-dipole (ZD) + Lennard-Jones nonbonded path and the SHAKE / leap-frog
 * step that consumes its forces.
 *
 * Conventions (all entry points):
 *   - Arrays are host pointers in md_run's column-major layout: cord(3,N) is
 *     C double[N][3] (xyz interleaved), grad(3,N,nfrg) is double[nfrg][N][3].
 *   - Atom indices in descriptors and outputs are 1-based (Fortran), converted
 *     at the boundary.
 *   - Return value: md_run's `ier` code for the routine the entry point
 *     replaces (0 ok; 28 cell overflow, 51 list overflow, 99 excluded pair not
 *     in the list, 1 SHAKE not converged, -2 SHAKE singular).  Values <= -100
 *     are runtime errors: JP_EINVAL / JP_ENOMEM / JP_ENODEV (-101..-103), or
 *     JP_EHIP_BASE - hipError_t = -1000 - hipError_t for a failed HIP call.
 *   - A context owns every device buffer and one HIP stream; the caller owns
 *     every host buffer.  One context per host thread; calls on one context
 *     are not re-entrant.
 *   - Units are md_run's: Angstrom, fs, amu, kcal/mol.
 */
#ifndef JPGPU_H
#define JPGPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JP_OK 0
#define JP_IER_CELL 28        /* celpar.f90:183-204 */
#define JP_IER_LIST 51        /* cutoff.f90:30-46 */
#define JP_IER_EXCL 99        /* cutoff.f90:90-96 */
#define JP_IER_SHAKE_CONV 1   /* shake.f90:286-319 */
#define JP_IER_SHAKE_SING (-2)/* shake.f90:266-284 */
#define JP_IER_STOP_MASS (-1)  /* remotv.f90:47-50: the stopped atoms have no mass */
#define JP_IER_STOP_SING (-3)  /* remotv.f90:86-88: md_run's ier is -2, as SHAKE's singular case */
#define JP_EINVAL (-101)
#define JP_ENOMEM (-102)
#define JP_ENODEV (-103)
#define JP_EHIP_BASE (-1000)  /* -1000 - hipError_t */

#define JP_PREC_FP64 0        /* plain FP64 pair arithmetic, FP64 accumulation */
#define JP_PREC_FP32_FIXED 1  /* FP32 pair arithmetic, 64-bit fixed-point accumulation */

/*
 * System description: the module state md_run builds in its INPUT phase and
 * in incel (inptpl.f90 types/LJ, inpcls.f90 chgmod/itab14/n_matrix,
 * setshk.f90 SHAKE groups, inpmd.f90 box and keywords, incel.f90 ZD
 * constants and the excess-pair list).
 */
typedef struct jp_sys_desc {
    int32_t natom;             /* ixnatm */
    int32_t ntyp;              /* iyntyp: number of LJ atom types */
    int32_t nfrag;             /* cluster classes (max icls) */
    int32_t nfrg;              /* energy classes (max n_matrix) */
    const double *x;           /* cord(3,N) initial coordinates */
    const double *mass;        /* fxmass(N) */
    const double *charge;      /* fxchrg(N), elementary charges */
    const int32_t *atype;      /* ixatyp(N), 1-based */
    const double *rmin2;       /* fyvwrd(ntyp): AMBER Rmin/2 */
    const double *eps;         /* fyvwme(ntyp) */
    const int32_t *icls;       /* icls(N), 1-based cluster class */
    const int32_t *n_matrix;   /* n_matrix(nfrag,nfrag), column-major, 1-based classes */
    int32_t nexcl;             /* unique excluded pairs (itab14 / 2) */
    const int32_t *excl;       /* [nexcl][2] 1-based */
    int32_t nexcess;           /* ZD excess pairs: bonds, angle 1-3, 1-4 (iytnbf=1) */
    const int32_t *excess;     /* [nexcess][2] 1-based, TPL order (incel.f90:665-688) */
    int32_t nshake;            /* SHAKE groups (iugshk) */
    const int32_t *shake_natom;/* [nshake] 2, 3 or 4 atoms (shak2e, shake3, shake4) */
    const int32_t *shake_atoms;/* [nshake][4] 1-based, unused slots 0 */
    const double *shake_dist2; /* [nshake][6] fudshk: squared targets, pair order 1-2,2-3,1-3,1-4,3-4,2-4 */
    double shake_tol;          /* fustol (COVSHK), md_run default 1e-6 */
    int32_t shake_maxit;       /* iuslop (LIMSHK), md_run default 1000 */
    int32_t cell_level;        /* CELLVL (nlev); 0 -> md_run default 3 */
    double celwal[6];          /* XMINCL XMAXCL YMINCL YMAXCL ZMINCL ZMAXCL */
    double cutoff;             /* CUTLEN (fycutl) */
    double diel;               /* DIEVAL (fydiel) */
    double dt;                 /* TIMEST (fs) */
    int32_t lmd;               /* 1: CLSTMETHOD=LMD, grad = l^2 g1 + l g2 + g3 (verlet.f90:67-91) */
    double lambda;             /* LAMBDA */
    /* Bonded terms of dnergv_PB (dnergv.f90:62-193): the bonds and angles left
     * after setshk removed the SHAKE-constrained ones (setshk.f90 reabnd /
     * reaang), i.e. md_run's iypbnd(:,1:nbnd) and iypang(:,1:nang) in its own
     * order.  nbond = nangle = 0 (water) means no bonded terms. */
    int32_t nbond;             /* nbnd = iynbnd - iyndbn */
    const int32_t *bond_atoms; /* [nbond][2] iypbnd, 1-based */
    const double *bond_fc;     /* fyfbnd: E = fc (r - r0)^2, kcal/mol/A^2 */
    const double *bond_r0;     /* fyqbnd, A */
    const int32_t *bond_cls;   /* bndcls, 1-based energy class; NULL: max(n_matrix) (SCALED_TERM default) */
    int32_t nangle;            /* nang = iynang - iyndag */
    const int32_t *angle_atoms;/* [nangle][3] iypang, 1-based */
    const double *angle_fc;    /* fyfang: E = fc (theta - theta0)^2, kcal/mol/rad^2 */
    const double *angle_theta0;/* fyqang, radians */
    const int32_t *angle_cls;  /* angcls, 1-based; NULL: max(n_matrix) */
    /* Torsions and improper torsions of dnei14_PB / dnetor_PB (dnergv.f90:197-429),
     * as incel leaves them (force constants below epsilon dropped,
     * incel.f90:993-1061): periodicity (1-6), phase in radians, division number. */
    int32_t ntors;             /* ntor */
    const int32_t *tors_atoms; /* [ntors][4] iyptor, 1-based */
    const double *tors_fc;     /* fyftor: E = fc (1 + cos(n phi + phase)) / div */
    const double *tors_rot;    /* fytrot (n) */
    const double *tors_phase;  /* fytphs */
    const int32_t *tors_div;   /* iytdiv */
    const int32_t *tors_cls;   /* torcls, 1-based; NULL: n_matrix of the two middle atoms (SCALED_TERM default) */
    int32_t nimp;              /* nimp: same layout (iypimp, fyfimp, fyirot, fyiphs, iyidiv) */
    const int32_t *imp_atoms;
    const double *imp_fc;
    const double *imp_rot;
    const double *imp_phase;
    const int32_t *imp_div;
    const int32_t *imp_cls;    /* impcls, 1-based; NULL: max(n_matrix) */
    /* 1-4 pairs of dne14_PB (dnergv.f90:433-516, incel.f90:975-984, 1084-1087) */
    int32_t n14;               /* n14int */
    const int32_t *p14_atoms;  /* [n14][2] iyp14, 1-based */
    const double *p14_vdw;     /* fytvws14: 1-4 LJ scale */
    const double *p14_elec;    /* fytess14 after incel: scale * ccoef * q_i * q_j */
    const int32_t *p14_cls;    /* i14cls, 1-based; NULL: n_matrix of the pair */
} jp_sys_desc;

/* Per-monitoring-step record of jp_run_nve (monmd.f90 / md.f90:618-628). */
typedef struct jp_energy_log {
    int32_t step;              /* idloop */
    double etot;               /* K + Epot (Epot lambda-scaled under LMD) */
    double epot;               /* potential energy */
    double ekin;               /* K from (v(t-dt/2)+v(t+dt/2))/2 (caltmp.f90) */
    double eelec[3];           /* ELE15 per energy class (enevcc(9,:)) */
    double evdw[3];            /* VDW15 per energy class (enevcc(8,:)) */
    double ebond[3];           /* BOND per energy class (enevcc(2,:)) */
    double eangle[3];          /* ANGLE per energy class (enevcc(3,:)) */
    double etors[3];           /* TORSION (enevcc(4,:)) */
    double eimp[3];            /* IMPROPER (enevcc(5,:)) */
    double evdw14[3];          /* 1-4 VDW (enevcc(6,:)) */
    double eelec14[3];         /* 1-4 ELEC (enevcc(7,:)) */
    double lambda;             /* lambda after the step (LMD: the first column of NAMETO,
                                  md.f90:578-583; moves under ALSD, jp_set_alsd) */
    double lambda_v;           /* its velocity after the step (ALSD; 0 otherwise) */
} jp_energy_log;

typedef struct jp_ctx jp_ctx;

/* incel + inpcls + setshk: build a context from a descriptor; uploads the
 * topology and the initial coordinates.  (incel.f90:2-689) */
int jp_ctx_create(const jp_sys_desc *desc, jp_ctx **out);
int jp_ctx_destroy(jp_ctx *ctx);

/* Host <-> device coordinates and velocities (cord(3,N), vel(3,N)). Either
 * pointer may be NULL. */
int jp_set_state(jp_ctx *ctx, const double *cord, const double *vel);
int jp_get_state(jp_ctx *ctx, double *cord, double *vel);

/* celprp: bin atoms into cells and build the pair list at the given
 * coordinates (NULL: the context's device coordinates).  (celprp.f90:2-175,
 * celpar.f90:2-244, cutoff.f90:2-505) */
int jp_build_list(jp_ctx *ctx, const double *cord);

/* Parity export of the last list: reference cell (mx,my,mz) of every atom
 * (celpar.f90:117-122 numbering), and the full (DOUB) neighbour list as CSR
 * over original atom order with ascending 1-based partner indices.  Pass
 * idx=NULL to query *npairs (sum of list lengths) only. */
int jp_get_list(jp_ctx *ctx, int32_t *cell_of_atom, int64_t *npairs, int32_t *ptr, int32_t *idx,
                int64_t cap);

/* cmmdo -> enear_ZD_*_PB plus the ZD self and excess terms: accumulates the
 * nonbonded gradient into grad(3,N,nfrg) and writes the per-class energies
 * (cmmdo.f90:2-78, enear.f90:1-593).  cord NULL: use device coordinates;
 * grad NULL: keep the gradient on the device. */
int jp_nonbond(jp_ctx *ctx, const double *cord, double *grad, double *eelec, double *evdw,
               double *ehyd, int precision);

/* dnergv_PB's bonded part: dnebnd_PB, dneang_PB and dnei14_PB (torsions,
 * impropers, 1-4 LJ and electrostatics), dnergv.f90:29-516.  Accumulates into
 * grad(3,N,nfrg) and writes eterm(6,nfrg) = enevcc(2:7,1:nfrg) (bond, angle,
 * torsion, improper, 1-4 vdW, 1-4 elec).  cord NULL: device coordinates;
 * grad NULL: energies only; eterm may be NULL. */
int jp_bonded(jp_ctx *ctx, const double *cord, double *grad, double *eterm);

/* shake_PB: constrain cord (x(t+dt)) against cord_old (x(t)) in place.
 * (shake.f90:2-755) */
int jp_shake(jp_ctx *ctx, double *cord, const double *cord_old);

/* runmd/verlet for METHOD=MICR: nsteps leap-frog steps on the device state
 * (pair forces, bonded terms, kick, drift, SHAKE, v=(x-x_old)/dt).  The
 * context counts steps like md_run's idloop (from 0, or from jp_set_step's
 * value, e.g. a restart's loop number): the list is rebuilt after every step
 * whose idloop is a multiple of `update` (md.f90:655-660), and energies are
 * recorded at every step whose idloop is a multiple of `nmon` (OUTLOG,
 * md.f90:604-633).  log has room for nsteps/nmon + 1 records (nmon > 0); the
 * recorded steps come first, in order, and the unused records get step = -1.
 * (md.f90:519-694, verlet.f90:2-226).  After jp_set_canonical the step is
 * verlethe's METHOD=CANO step instead (verlethe.f90:79-757). */
int jp_run_nve(jp_ctx *ctx, int32_t nsteps, int32_t update, int32_t nmon, int32_t precision,
               jp_energy_log *log);

/* Captures, ahead of time, the HIP graphs of plain steps that a following
 * jp_run_nve with the same arguments will replay from the current step
 * counter, so that the run itself captures none (unless a list outgrows its
 * segments).  No step is run.  A preparation helper with no counterpart in
 * the reference. */
int jp_prepare_graphs(jp_ctx *ctx, int32_t nsteps, int32_t update, int32_t nmon, int32_t precision);

/* md_run's step counter idloop: inirst sets it from the restart file's loop
 * number (inprst call, md.f90:1132), so that list rebuilds and OUTLOG steps of a
 * continued run fall where md_run's do.  A new context starts at 0. */
int jp_set_step(jp_ctx *ctx, int64_t idloop);
int jp_get_step(jp_ctx *ctx, int64_t *idloop);

/* md_run's restart set-up (inirst, md.f90:1104-1263): one step from the
 * current state, recorded in log[0] (NULL for none; one record per system,
 * step = the current idloop) as md_run's energy-continuity check, after
 * which the coordinates are put back and the velocities become the ones the
 * step's last verlethe substep started from (md.f90:1240-1250 restores vel
 * from verlethe's velback).  Without a SHAKE retry these are the restart's
 * own velocities; after one (METHOD= CANO) they are those at the start of
 * the step's last substep.  The step counter does not advance.  Returns the
 * step's SHAKE ier, with the state restored all the same. */
int jp_restart_step(jp_ctx *ctx, int32_t precision, jp_energy_log *log);

/* Seconds spent in the last jp_run_nve, measured with HIP events on the
 * context stream: total, pair kernel, list build, the rest (SHAKE,
 * integration, launch gaps).  The pair time is the mean over the pair
 * launches issued from the host (energy steps, or every step with
 * JP_OPT_GRAPHS = 0) times the number of steps; steps replayed from a HIP
 * graph are not timed one by one. */
int jp_last_timing(jp_ctx *ctx, double *t_total, double *t_pair, double *t_list, double *t_step);

/* Canonical MD (md_run METHOD= CANO, verlethe.f90:79-757 with constant
 * lambda, STOPCE= NO and HEATLO= 0) for the following jp_run_nve calls:
 * every step scales the unconstrained half-step velocities by
 * chi = sqrt(temps / T) (caltmp.f90), then runs up to icslop rounds of SHAKE
 * and temperature (md_run's defaults 3 and fcstol 1e-3, main.f90:166-167)
 * until every temperature group is within fcstol * temps.  group: NULL for
 * one temperature per system (TEMPCO= NO), or per atom a positive label of
 * its temperature group (TEMPCO= YES: md_run's ixamol2, md.f90:136-152);
 * in a batch, labels are per system and the array covers all atoms in order.
 * Degrees of freedom per group: 3 per atom minus the constraints of the SHAKE
 * groups whose first atom is in it.  on = 0 returns to METHOD= MICR.  For a
 * single system, a step whose SHAKE fails is redone as md_run's verlethe
 * does (verlethe.f90:552-573, 745-752): undone, then substeps of half the
 * length (halving again on a failure, doubling back after two successes)
 * until they add up to dt; a substep below dt/100 ends the run with the
 * SHAKE ier.  In a batch a SHAKE failure ends the run with its ier.  Returns JP_EINVAL for temps <= 0, icslop < 1, a
 * label < 1 or a group without degrees of freedom. */
int jp_set_canonical(jp_ctx *ctx, int32_t on, double temps, int32_t icslop, double fcstol, const int32_t *group);

/* md_run's heating loop (HEATLO, STARTT; verlethe.f90:419-423, 455-459,
 * 600-604, 641-646, 671-675) for the canonical runs that follow: while
 * md_run's step count idloop (jp_set_step, the restart's loop count) is at
 * most heatlo, the target temperature of the step is
 * startt + ((temps - startt) / heatlo) idloop in place of temps (the
 * velocity scaling and the end-of-round test; without SHAKE groups the
 * current temperature is still taken as temps), the restart set-up step
 * (jp_restart_step) using the restart's idloop.  Those steps run as host
 * launches, the later ones as step graphs again.  heatlo = 0 (md_run's
 * default) switches it off.  Returns JP_EINVAL for heatlo < 0 or startt < 0. */
int jp_set_heating(jp_ctx *ctx, int32_t heatlo, double startt);

/* Centre-of-mass stopping (md_run STOPCE= TRAN, ROTA, BOTH; remotv.f90:2-131)
 * for the steps that follow: mode JP_STOP_TRAN takes the first natom atoms'
 * centre-of-mass velocity out of their velocities, JP_STOP_ROTA their
 * rotation about the centre of mass, JP_STOP_BOTH both, JP_STOP_NO nothing;
 * natom is md_run's ixcend(NSTPCN), the atoms of the first NSTPCN chains
 * (molecule copies).  It acts where md_run calls remotv: in METHOD= MICR
 * between the kick and the drift (verlet.f90:156-169), in the canonical
 * step after the velocity update of every round (verlethe.f90:489-502,
 * 705-718), and in jp_start_step, always about the coordinates at the start
 * of the step.  jp_set_canonical, called after it, takes md_run's 3 (TRAN,
 * ROTA) or 6 (BOTH) degrees of freedom off (md.f90:114-121; under
 * temperature groups, md.f90:183-215, off the groups whose labels the first
 * nkind atoms carry, nkind being the atoms from atom 1 of atom 1's molecule
 * kind, in proportion to their degrees of freedom when there are several).
 * A run whose stopped atoms have no mass ends with JP_IER_STOP_MASS, one
 * whose inertia tensor is singular with JP_IER_STOP_SING.  md_run's own
 * STOPCE= TRAN reads an unset angular velocity (remotv.f90:122-127; the
 * angular velocity is set only for ROTA and BOTH at 106-116); here it is
 * zero.  Returns JP_EINVAL for another mode, natom outside 1..N, nkind
 * outside 0..N, or a call after jp_set_canonical. */
#define JP_STOP_NO 0
#define JP_STOP_TRAN 1
#define JP_STOP_ROTA 2
#define JP_STOP_BOTH 3
int jp_set_stopce(jp_ctx *ctx, int32_t mode, int32_t natom, int32_t nkind);

/* md_run's initial step of a run without a restart (RESTAR= NO; inista,
 * md.f90:1056-1065, staver.f90:27-207), from the context's coordinates and
 * velocities (zero under INITIA= ZERO): the pair list of x, the forces at x
 * (the lambda-combined gradient, without the METHOD= GE weight; lambda does
 * not move), v += (-dt c / 2) g / m, STOPCE (jp_set_stopce) about x,
 * x += dt v and SHAKE of x; v keeps the unconstrained half-kick velocity, as
 * md_run's does.  The step count stays (0 for md_run's fresh start), and
 * the first jp_run_nve step follows.  Returns ier (SHAKE's, or remotv's). */
int jp_start_step(jp_ctx *ctx, int32_t precision);

/* Multicanonical MD (md_run METHOD= GE, idfvsl = 4, verlethe.f90:196-233) for
 * the following jp_run_nve calls, on top of jp_set_canonical's temperature
 * control: every step's gradient is multiplied by the weight
 * (gascst/joucal) 1e-3 tempce dprobl of the step's potential energy E
 * (epotcc(1,1)), with dprobl = alphalw for E <= celw, alphaup for E >= ceup,
 * else the polynomial c[i*(ndeg+1) + j] (r - low[i])^j, j = 0..ndeg, of the
 * first window i with r = (E - celw)/(ceup - celw) < high[i] (the last
 * window beyond it); the parameters of the NAMEGE= file (inmcmd.f90:20-31).
 * Every step then evaluates energies, and E never leaves the device.  One
 * system with one energy class (no LMD: ALSD is jp_set_alsd; two-class McMD
 * is not provided).  on = 0 switches it off; on = 1 switches ALSD off.
 * Returns JP_EINVAL for other systems, nwindow < 1, ndeg < 0, celw >= ceup
 * or non-finite parameters; jp_run_nve returns JP_EINVAL while it is on
 * without jp_set_canonical. */
int jp_set_mcmd(jp_ctx *ctx, int32_t on, int32_t nwindow, int32_t ndeg, const double *c, const double *low,
                const double *high, double alphalw, double alphaup, double celw, double ceup, double tempce);

/* Adaptive lambda square dynamics (md_run METHOD= GE under GEPSME= LMD:
 * idfvsl = 4 with cluster_method LMD, verlethe.f90:235-290) for the
 * following jp_run_nve calls, on top of jp_set_canonical's temperature
 * control (call it after jp_set_canonical; a new jp_set_canonical turns it
 * off).  lambda becomes a dynamical variable of mass lambda_m (LWEIGH;
 * <= 0: md_run's default, the class-1 atoms' mass added in atom order,
 * inpcls.f90:348-357) driven by
 *   lambda_f = -2 E1 lambda - E2 - (gascst/joucal) 1e-3 tempce dprobl,
 * E1 and E2 the step's unscaled class-1 and class-2 potential energies and
 * dprobl the polynomial c[i*(ndeg+1) + j] (lambda - low[i])^j of the first
 * window i with lambda < high[i] (the last one beyond it), plus
 * lower * (celw - lambda) below celw and upper * (lambda - ceup) above ceup
 * (the NAMEGE= file of inmcmd.f90:20-41).  temp_control is LTCNTL
 * (l_temp_control): lambda's velocity is scaled with its temperature group
 * and its degree of freedom is counted; its kinetic term m_l lv^2 is in the
 * temperature of the one group always (caltmp), of the group with the most
 * degrees of freedom under temperature groups only with temp_control
 * (caltmp2, md.f90:221-228).  lambda and its velocity stay on the device
 * during a run (jp_get_lambda / jp_set_lambda, restart files:
 * outrst.f90:61); every step evaluates energies.  Systems need LMD with
 * three energy classes.  on = 0 switches it off (and jp_set_mcmd).  Returns
 * JP_EINVAL without METHOD= CANO, for other systems, nwindow < 1, ndeg < 0
 * or non-finite parameters. */
int jp_set_alsd(jp_ctx *ctx, int32_t on, int32_t nwindow, int32_t ndeg, const double *c, const double *low,
                const double *high, double lower, double upper, double celw, double ceup, double tempce,
                double lambda_m, int32_t temp_control);
/* lambda and its velocity (COMCMM's lambda, lambda_v; the restart file's
 * last record, inprst.f90:56-66, outrst.f90:61).  Setting them also sets the
 * class factors of the following evaluations. */
int jp_get_lambda(jp_ctx *ctx, double *lambda, double *lambda_v);
int jp_set_lambda(jp_ctx *ctx, double lambda, double lambda_v);

/* The cell multipole method in place of ZD (md_run CAL15M= CMM with BOUNDA=
 * NO, cmmdo.f90:79-288, neartb.f90, celpol.f90) for the following
 * evaluations and runs of a non-periodic system: cminsiz is CELSIZ (A, md_run
 * default 6), nlev CELLVL (default 3; raised, as incel.f90:42-52 does, until
 * the level-1 cells cover the extent of the current coordinates), cmmupd
 * CMMUPD (default 5): the far field is recomputed on the first step after
 * every cmmupd-th and reused in between (md.f90:679-684), the tables and
 * multipoles at every rebuild (jp_build_list, jp_run_nve's update).  The
 * near field is Coulomb over every pair in a cell and its 26 neighbours
 * without cutoff, LJ where both atoms carry it, plus LJ of the cells out to
 * ceil(cutoff / cminsiz) for pairs within the cutoff at the rebuild; there is
 * no ZD self or excess term.  Everything runs in FP64 whatever precision a
 * call asks for.  jp_build_list returns md_run's ier 28 (an atom off the
 * cells, too many in one), 51 (a table over its capacity) or 99 (an
 * excluded pair outside the tables); jp_nonbond computes the far field
 * afresh; jp_get_list returns the level-1 cell (1-based per axis) of every
 * atom and no pair list; jp_count_pairs returns JP_EINVAL.  A new jp_set_state
 * starts with the far field due.  on = 0 returns to ZD.  Returns JP_EINVAL
 * for a batch or a context of several systems, cminsiz <= 0 or below
 * cutoff / 15, nlev outside
 * 1..5, cmmupd < 1, or coordinates that need more than 5 levels (md_run
 * stops there). */
int jp_set_cmm(jp_ctx *ctx, int32_t on, double cminsiz, int32_t nlev, int32_t cmmupd);
/* Under CMM: mean durations in seconds of the near-field kernels (the
 * 27-cell Coulomb/LJ pass and the second-nearest LJ pass) and of the
 * far-field kernel, over reps launches of each back to back at the device
 * coordinates, between HIP events.  A measurement helper with no counterpart
 * in the reference; the positions, velocities and the stored far field are
 * those of a fresh evaluation afterwards. */
int jp_time_cmm(jp_ctx *ctx, int32_t reps, double *near_s, double *far_s);

/* Mean duration in seconds of the force-step pair kernel (the kernel of every
 * step without energies) over reps launches back to back on the current list
 * and coordinates, measured with two HIP events around the launches.  A
 * measurement helper with no counterpart in the reference; it leaves the
 * positions and velocities alone. */
int jp_time_pair(jp_ctx *ctx, int32_t precision, int32_t reps, double *seconds);

/* Unique pairs with r <= cutoff at the device coordinates (the
 * pair-interactions/s numerator), those of them where both atoms carry LJ,
 * and the unique pairs in the list (r <= cutoff+1.5 at the last rebuild).
 * nlj and nlisted may be NULL. */
int jp_count_pairs(jp_ctx *ctx, int64_t *npairs, int64_t *nlj, int64_t *nlisted);

/* jp_count_pairs for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  It calls no allocating routine: its
 * per-cluster partial counts live in a fixed scratch array the library holds
 * on every GPU it runs on (6 MB, up to 262,144 i-clusters of 8 atoms, i.e.
 * about 2M atoms); a larger system returns JP_EINVAL.  Results equal
 * jp_count_pairs's.  The scratch is shared by every call on one GPU: calls
 * that overlap on one GPU (several host threads, several contexts) race. */
int jp_count_pairs_noalloc(jp_ctx *ctx, int64_t *npairs, int64_t *nlj, int64_t *nlisted);

/*
 * Batched ensemble (replaces j_presto_mpi / j_presto_master: N independent
 * md_run jobs, j_presto_mpi:27-44): every system of the batch shares each
 * kernel launch.
 */
typedef struct jp_batch jp_batch;
int jp_batch_create(const jp_sys_desc *descs, int32_t nsys, jp_batch **out);
int jp_batch_run_nve(jp_batch *b, int32_t nsteps, int32_t update, int32_t nmon, int32_t precision,
                     jp_energy_log *log /* [nsys][nsteps/nmon + 1] */);
int jp_batch_set_step(jp_batch *b, int64_t idloop);   /* jp_set_step for the batch's shared counter */
int jp_batch_restart_step(jp_batch *b, int32_t precision, jp_energy_log *log /* [nsys] */);
int jp_batch_prepare_graphs(jp_batch *b, int32_t nsteps, int32_t update, int32_t nmon, int32_t precision);
int jp_batch_set_state(jp_batch *b, int32_t isys, const double *cord, const double *vel);
int jp_batch_get_state(jp_batch *b, int32_t isys, double *cord, double *vel);
int jp_batch_last_timing(jp_batch *b, double *t_total);
int jp_batch_destroy(jp_batch *b);

/* Run options (no counterpart in md_run, except that the autotuner plays the
 * role of SIMD_check's runtime choice of i_vec, SIMD_check.f90:2-42).
 *   JP_OPT_GRAPHS      1 (default): the plain steps between list rebuilds and
 *                      energy steps of jp_run_nve / jp_batch_run_nve run as
 *                      captured HIP graphs; 0: every kernel launched from the
 *                      host.  Results are identical either way.
 *   JP_OPT_AUTOTUNE    1: the next run tunes the box-grid cell size (the
 *                      4.5-9 A grid and the grids with cells half and twice as
 *                      wide), the list kernel's wavefronts per block (1/2/4/8)
 *                      and tile width (1/2/4 batches of 8 j-clusters per pass),
 *                      and the FP32 pair kernel's wavefronts per i-cluster
 *                      (2/4/8), timing each on the live list, and keeps the
 *                      grid with the least rebuild time / update + pair time
 *                      with its fastest shapes; 0 (default): fixed choices.
 *   JP_OPT_HALF_LIST   1: the FP32/fixed-point path runs on a half list (each unique
 *                      pair once, its force applied to both atoms, the j side as
 *                      one 64-bit fixed-point atomic add per entry and component);
 *                      0 (default): md_run's full list, each atom summing its own
 *                      pairs.  Results agree within the FP32 tolerance.  On MI355X
 *                      the half list is the slower one: its 3 atomics per list
 *                      entry cost more than the pair work they save.  A system
 *                      whose md_run cell table drops pairs asymmetrically keeps the
 *                      full list; the FP64 path always uses it; jp_get_list exports
 *                      the full list either way.  A run or call whose precision
 *                      needs the other kind of list builds it at the current
 *                      coordinates.
 *   JP_OPT_PAIR_WAVES  wavefronts per i-cluster in the FP32 force-step pair kernel:
 *                      full list 2, 4 (default), 8 (1 counts as 2); half list 1,
 *                      2 (default), 4, 8; energy steps and the FP64 path keep
 *                      their defaults.
 *                      It regroups FP32 partial sums: FP32/fixed-point results are
 *                      deterministic for a given value and agree within the FP32
 *                      tolerance across values.
 *   JP_OPT_LIST_WAVES  i-clusters per block in the list kernel (1, 2, 4 = default, 8);
 *                      the list does not depend on it.
 *   JP_OPT_LIST_BATCH  batches of 8 j-clusters the list kernel tests per pass
 *                      (1, 2 = default, 4); the list does not depend on it.
 *   JP_OPT_GRID_SHIFT  box-grid cell size: 0 (default) cells of 4.5-9 A, k > 0
 *                      2^k times narrower, k < 0 2^-k times wider (-3..3; at most
 *                      512 and at least 2 cells per edge).  The pair set does not
 *                      depend on it; the slot order does, so it regroups partial
 *                      sums like JP_OPT_PAIR_WAVES.  The next evaluation rebuilds
 *                      the list.
 *   JP_OPT_LIST_SEGMENT  list entries per i-cluster segment at the next rebuild
 *                      only (a multiple of 64; set, not read); the library grows
 *                      segments by itself and rebuilds when one overflows, so this
 *                      only exercises that path (results do not depend on it).
 *   JP_OPT_CELL_TABLE  1 (default): the pair list holds the pairs md_run's own
 *                      periodic neighbour-cell table lets it test (incel.f90:
 *                      456-508), bit for bit like cutatm_PB.  That table wraps
 *                      a cell range that can be up to two cells longer than the
 *                      box edge, and for some edges (40.75-47 A, 54.25-60.5 A,
 *                      ..., 94.5-99.75 A at a 12 A cutoff) md_run never tests
 *                      pairs across the periodic boundary that lie inside the
 *                      cutoff.  0: every pair within rlim by the minimum image
 *                      (the complete list).  Where md_run's table is undefined
 *                      (its cell range leaves 1..nv, e.g. a 216 A box at
 *                      CELLVL 3) the list is the complete one either way.  A
 *                      change takes effect at the next list build.
 * Setting a shape turns the autotuner off; jp_get_option reads the current
 * (possibly tuned) values. */
#define JP_OPT_GRAPHS 1
#define JP_OPT_AUTOTUNE 2
#define JP_OPT_PAIR_WAVES 3
#define JP_OPT_LIST_WAVES 4
#define JP_OPT_LIST_SEGMENT 5
#define JP_OPT_CELL_TABLE 6
#define JP_OPT_GRAPH_CAPTURES 7   /* read only: step graphs captured during the last run */
#define JP_OPT_HALF_LIST 8
#define JP_OPT_LIST_BATCH 9
#define JP_OPT_GRID_SHIFT 10
int jp_set_option(jp_ctx *ctx, int32_t option, int32_t value);
int jp_batch_set_option(jp_batch *b, int32_t option, int32_t value);
int jp_get_option(jp_ctx *ctx, int32_t option, int32_t *value);
int jp_batch_get_option(jp_batch *b, int32_t option, int32_t *value);

/* jp_set_canonical for every system of a batch (group: labels for all atoms
 * of the batch in order, or NULL). */
int jp_batch_set_canonical(jp_batch *b, int32_t on, double temps, int32_t icslop, double fcstol,
                           const int32_t *group);
/* jp_set_heating for every system of a batch. */
int jp_batch_set_heating(jp_batch *b, int32_t heatlo, double startt);
/* jp_set_stopce for every system of a batch: natom[s] and nkind[s] (nkind
 * may be NULL) per system in batch order. */
int jp_batch_set_stopce(jp_batch *b, int32_t mode, const int32_t *natom, const int32_t *nkind);
/* jp_start_step for every system of a batch. */
int jp_batch_start_step(jp_batch *b, int32_t precision);
/* jp_set_alsd, jp_get_lambda and jp_set_lambda for the systems of a batch
 * (one parameter set; lambda, its velocity and, with lambda_m <= 0, its
 * mass per system). */
int jp_batch_set_alsd(jp_batch *b, int32_t on, int32_t nwindow, int32_t ndeg, const double *c, const double *low,
                      const double *high, double lower, double upper, double celw, double ceup, double tempce,
                      double lambda_m, int32_t temp_control);
int jp_batch_get_lambda(jp_batch *b, int32_t isys, double *lambda, double *lambda_v);
int jp_batch_set_lambda(jp_batch *b, int32_t isys, double lambda, double lambda_v);


/* Library version and device name (for logs). */
const char *jp_version(void);

#ifdef __cplusplus
}
#endif
#endif
