! ISO_C_BINDING interface of libjpgpu (include/jpgpu.h): the MI355X path for
! md_run's ZD + LJ nonbonded forces, the pair list that feeds them, and the
! SHAKE / leap-frog step.
!
! Every routine returns md_run's ier code for the routine it replaces
! (0 ok; 28 cell overflow, 51 list overflow, 99 excluded pair missing from
! the list, 1 SHAKE not converged, -2 SHAKE singular); values <= -100 are
! runtime errors (-101..-103: JP_EINVAL, JP_ENOMEM, JP_ENODEV; -1000 - e for
! a HIP call that failed with hipError_t e).  Arrays are passed in md_run's own layout: cord(3,N),
! vel(3,N), grad(3,N,nfrg), 1-based atom indices.  The reference routines
! each entry point stands in for:
!   jp_ctx_create   incel.f90:2 + inpcls.f90 + setshk.f90 (set-up)
!   jp_build_list   celprp.f90:2 (celpar_PB, cutatm_PB, cutoff_PB)
!   jp_nonbond      cmmdo.f90:2 -> enear_ZD_*_PB (enear.f90:1-593)
!   jp_bonded       dnergv.f90:29-516 (dnebnd_PB, dneang_PB, dnei14_PB)
!   jp_shake        shake.f90:2 (shake_PB)
!   jp_run_nve      md.f90:519-694 + verlet.f90:2 (METHOD= MICR loop)
!   jp_batch_*      j_presto_mpi:27-44 (independent runs, one launch)
module jpgpu
  use iso_c_binding
  implicit none
  private :: c_strlen

  integer(c_int), parameter :: JP_OK = 0
  integer(c_int), parameter :: JP_IER_CELL = 28, JP_IER_LIST = 51, JP_IER_EXCL = 99
  integer(c_int), parameter :: JP_IER_SHAKE_CONV = 1, JP_IER_SHAKE_SING = -2
  integer(c_int), parameter :: JP_PREC_FP64 = 0, JP_PREC_FP32_FIXED = 1
  integer(c_int32_t), parameter :: JP_OPT_GRAPHS = 1, JP_OPT_AUTOTUNE = 2
  integer(c_int32_t), parameter :: JP_OPT_PAIR_WAVES = 3, JP_OPT_LIST_WAVES = 4, JP_OPT_LIST_SEGMENT = 5
  integer(c_int32_t), parameter :: JP_OPT_CELL_TABLE = 6, JP_OPT_HALF_LIST = 8, JP_OPT_LIST_BATCH = 9
  integer(c_int32_t), parameter :: JP_OPT_GRID_SHIFT = 10

  ! System description: the module state md_run builds in its input phase.
  ! Same field order and C layout as jp_sys_desc in include/jpgpu.h.
  type, bind(C) :: jp_sys_desc
     integer(c_int32_t) :: natom = 0, ntyp = 0, nfrag = 1, nfrg = 1
     type(c_ptr) :: x = c_null_ptr, mass = c_null_ptr, charge = c_null_ptr
     type(c_ptr) :: atype = c_null_ptr, rmin2 = c_null_ptr, eps = c_null_ptr
     type(c_ptr) :: icls = c_null_ptr, n_matrix = c_null_ptr
     integer(c_int32_t) :: nexcl = 0
     type(c_ptr) :: excl = c_null_ptr
     integer(c_int32_t) :: nexcess = 0
     type(c_ptr) :: excess = c_null_ptr
     integer(c_int32_t) :: nshake = 0
     type(c_ptr) :: shake_natom = c_null_ptr, shake_atoms = c_null_ptr, shake_dist2 = c_null_ptr
     real(c_double) :: shake_tol = 1.0d-6
     integer(c_int32_t) :: shake_maxit = 1000
     integer(c_int32_t) :: cell_level = 0
     real(c_double) :: celwal(6) = 0.0d0
     real(c_double) :: cutoff = 12.0d0, diel = 1.0d0, dt = 2.0d0
     integer(c_int32_t) :: lmd = 0
     real(c_double) :: lambda = 0.0d0
     ! bonds and angles left after setshk (iypbnd(:,1:nbnd), iypang(:,1:nang))
     integer(c_int32_t) :: nbond = 0
     type(c_ptr) :: bond_atoms = c_null_ptr, bond_fc = c_null_ptr, bond_r0 = c_null_ptr
     type(c_ptr) :: bond_cls = c_null_ptr
     integer(c_int32_t) :: nangle = 0
     type(c_ptr) :: angle_atoms = c_null_ptr, angle_fc = c_null_ptr, angle_theta0 = c_null_ptr
     type(c_ptr) :: angle_cls = c_null_ptr
     ! torsions and impropers as incel leaves them, 1-4 pairs (iyp14, fytvws14, fytess14, i14cls)
     integer(c_int32_t) :: ntors = 0
     type(c_ptr) :: tors_atoms = c_null_ptr, tors_fc = c_null_ptr, tors_rot = c_null_ptr
     type(c_ptr) :: tors_phase = c_null_ptr, tors_div = c_null_ptr, tors_cls = c_null_ptr
     integer(c_int32_t) :: nimp = 0
     type(c_ptr) :: imp_atoms = c_null_ptr, imp_fc = c_null_ptr, imp_rot = c_null_ptr
     type(c_ptr) :: imp_phase = c_null_ptr, imp_div = c_null_ptr, imp_cls = c_null_ptr
     integer(c_int32_t) :: n14 = 0
     type(c_ptr) :: p14_atoms = c_null_ptr, p14_vdw = c_null_ptr, p14_elec = c_null_ptr
     type(c_ptr) :: p14_cls = c_null_ptr
  end type jp_sys_desc

  ! One monitoring record of jp_run_nve (md.f90:618-628).
  type, bind(C) :: jp_energy_log
     integer(c_int32_t) :: step
     real(c_double) :: etot, epot, ekin
     real(c_double) :: eelec(3), evdw(3)
     real(c_double) :: ebond(3), eangle(3), etors(3), eimp(3), evdw14(3), eelec14(3)
     real(c_double) :: lambda, lambda_v
  end type jp_energy_log

  interface
     integer(c_int) function jp_ctx_create(desc, ctx) bind(C, name='jp_ctx_create')
       import :: c_int, c_ptr, jp_sys_desc
       type(jp_sys_desc), intent(in) :: desc
       type(c_ptr), intent(out) :: ctx
     end function jp_ctx_create

     integer(c_int) function jp_ctx_destroy(ctx) bind(C, name='jp_ctx_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function jp_ctx_destroy

     ! cord / vel: c_loc of real(8) (3,N) arrays, or c_null_ptr
     integer(c_int) function jp_set_state(ctx, cord, vel) bind(C, name='jp_set_state')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, cord, vel
     end function jp_set_state

     integer(c_int) function jp_get_state(ctx, cord, vel) bind(C, name='jp_get_state')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, cord, vel
     end function jp_get_state

     integer(c_int) function jp_build_list(ctx, cord) bind(C, name='jp_build_list')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, cord
     end function jp_build_list

     integer(c_int) function jp_get_list(ctx, cell_of_atom, npairs, ptr, idx, cap) bind(C, name='jp_get_list')
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: ctx, cell_of_atom
       integer(c_int64_t), intent(out) :: npairs
       type(c_ptr), value :: ptr, idx
       integer(c_int64_t), value :: cap
     end function jp_get_list

     ! grad is accumulated (md_run zeroes it in dnergy before cmmdo)
     integer(c_int) function jp_nonbond(ctx, cord, grad, eelec, evdw, ehyd, precision) bind(C, name='jp_nonbond')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, cord, grad, eelec, evdw, ehyd
       integer(c_int), value :: precision
     end function jp_nonbond

     ! dnergv_PB's bonded part: grad accumulated, eterm(6,nfrg) = enevcc(2:7,:)
     integer(c_int) function jp_bonded(ctx, cord, grad, eterm) bind(C, name='jp_bonded')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, cord, grad, eterm
     end function jp_bonded

     integer(c_int) function jp_shake(ctx, cord, cord_old) bind(C, name='jp_shake')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, cord, cord_old
     end function jp_shake

     integer(c_int) function jp_run_nve(ctx, nsteps, update, nmon, precision, log) bind(C, name='jp_run_nve')
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: nsteps, update, nmon, precision
       type(c_ptr), value :: log
     end function jp_run_nve

     integer(c_int) function jp_last_timing(ctx, t_total, t_pair, t_list, t_step) bind(C, name='jp_last_timing')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: t_total, t_pair, t_list, t_step
     end function jp_last_timing

     ! md_run's step counter idloop (inirst sets it from the restart, md.f90:1132)
     integer(c_int) function jp_set_step(ctx, idloop) bind(C, name='jp_set_step')
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: idloop
     end function jp_set_step

     integer(c_int) function jp_get_step(ctx, idloop) bind(C, name='jp_get_step')
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), intent(out) :: idloop
     end function jp_get_step

     ! md_run's restart set-up step (inirst, md.f90:1104-1263): one step,
     ! then the coordinates back and the velocities of its last substep start
     integer(c_int) function jp_restart_step(ctx, precision, log) bind(C, name='jp_restart_step')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: precision
       type(c_ptr), value :: log
     end function jp_restart_step

     integer(c_int) function jp_count_pairs(ctx, npairs, nlj, nlisted) bind(C, name='jp_count_pairs')
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), intent(out) :: npairs, nlj, nlisted
     end function jp_count_pairs

     integer(c_int) function jp_set_canonical(ctx, on, temps, icslop, fcstol, group) &
          bind(C, name='jp_set_canonical')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx, group
       integer(c_int32_t), value :: on, icslop
       real(c_double), value :: temps, fcstol
     end function jp_set_canonical

     integer(c_int) function jp_set_mcmd(ctx, on, nwindow, ndeg, c, low, high, alphalw, alphaup, celw, ceup, &
          tempce) bind(C, name='jp_set_mcmd')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx, c, low, high
       integer(c_int32_t), value :: on, nwindow, ndeg
       real(c_double), value :: alphalw, alphaup, celw, ceup, tempce
     end function jp_set_mcmd

     integer(c_int) function jp_set_alsd(ctx, on, nwindow, ndeg, c, low, high, lower, upper, celw, ceup, &
          tempce, lambda_m, temp_control) bind(C, name='jp_set_alsd')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx, c, low, high
       integer(c_int32_t), value :: on, nwindow, ndeg, temp_control
       real(c_double), value :: lower, upper, celw, ceup, tempce, lambda_m
     end function jp_set_alsd

     integer(c_int) function jp_set_heating(ctx, heatlo, startt) bind(C, name='jp_set_heating')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: heatlo
       real(c_double), value :: startt
     end function jp_set_heating

     integer(c_int) function jp_set_stopce(ctx, mode, natom, nkind) bind(C, name='jp_set_stopce')
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: mode, natom, nkind
     end function jp_set_stopce

     integer(c_int) function jp_start_step(ctx, precision) bind(C, name='jp_start_step')
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: precision
     end function jp_start_step

     integer(c_int) function jp_set_cmm(ctx, on, cminsiz, nlev, cmmupd) bind(C, name='jp_set_cmm')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: on, nlev, cmmupd
       real(c_double), value :: cminsiz
     end function jp_set_cmm

     integer(c_int) function jp_get_lambda(ctx, lambda, lambda_v) bind(C, name='jp_get_lambda')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: lambda, lambda_v
     end function jp_get_lambda

     integer(c_int) function jp_set_lambda(ctx, lambda, lambda_v) bind(C, name='jp_set_lambda')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: lambda, lambda_v
     end function jp_set_lambda

     integer(c_int) function jp_batch_set_canonical(b, on, temps, icslop, fcstol, group) &
          bind(C, name='jp_batch_set_canonical')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: b, group
       integer(c_int32_t), value :: on, icslop
       real(c_double), value :: temps, fcstol
     end function jp_batch_set_canonical

     integer(c_int) function jp_time_pair(ctx, precision, reps, seconds) bind(C, name='jp_time_pair')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: precision, reps
       real(c_double), intent(out) :: seconds
     end function jp_time_pair

     integer(c_int) function jp_batch_create(descs, nsys, b) bind(C, name='jp_batch_create')
       import :: c_int, c_int32_t, c_ptr, jp_sys_desc
       type(jp_sys_desc), intent(in) :: descs(*)
       integer(c_int32_t), value :: nsys
       type(c_ptr), intent(out) :: b
     end function jp_batch_create

     integer(c_int) function jp_batch_run_nve(b, nsteps, update, nmon, precision, log) bind(C, name='jp_batch_run_nve')
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: b
       integer(c_int32_t), value :: nsteps, update, nmon, precision
       type(c_ptr), value :: log
     end function jp_batch_run_nve

     integer(c_int) function jp_batch_set_state(b, isys, cord, vel) bind(C, name='jp_batch_set_state')
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: b
       integer(c_int32_t), value :: isys
       type(c_ptr), value :: cord, vel
     end function jp_batch_set_state

     integer(c_int) function jp_batch_get_state(b, isys, cord, vel) bind(C, name='jp_batch_get_state')
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: b
       integer(c_int32_t), value :: isys
       type(c_ptr), value :: cord, vel
     end function jp_batch_get_state

     integer(c_int) function jp_batch_last_timing(b, t_total) bind(C, name='jp_batch_last_timing')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: b
       real(c_double), intent(out) :: t_total
     end function jp_batch_last_timing

     integer(c_int) function jp_batch_destroy(b) bind(C, name='jp_batch_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: b
     end function jp_batch_destroy

     integer(c_int) function jp_set_option(ctx, option, value) bind(C, name='jp_set_option')
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: option, value
     end function jp_set_option

     integer(c_int) function jp_batch_set_option(b, option, value) bind(C, name='jp_batch_set_option')
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: b
       integer(c_int32_t), value :: option, value
     end function jp_batch_set_option

     integer(c_int) function jp_get_option(ctx, option, value) bind(C, name='jp_get_option')
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: option
       integer(c_int32_t), intent(out) :: value
     end function jp_get_option

     integer(c_int) function jp_batch_get_option(b, option, value) bind(C, name='jp_batch_get_option')
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: b
       integer(c_int32_t), value :: option
       integer(c_int32_t), intent(out) :: value
     end function jp_batch_get_option

     type(c_ptr) function jp_version_c() bind(C, name='jp_version')
       import :: c_ptr
     end function jp_version_c

     integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
       import :: c_ptr, c_size_t
       type(c_ptr), value :: s
     end function c_strlen
  end interface

contains

  ! Library version string as a Fortran character value.
  function jp_version() result(v)
    character(len=:), allocatable :: v
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: i, n
    p = jp_version_c()
    n = int(c_strlen(p))
    call c_f_pointer(p, s, [n])
    allocate(character(len=n) :: v)
    do i = 1, n
       v(i:i) = s(i)
    end do
  end function jp_version

end module jpgpu
