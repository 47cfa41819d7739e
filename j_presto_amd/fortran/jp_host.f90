! Fortran host of libjpgpu.  Reads a system file, then does through the C ABI
! what md_run does through its own routines:
!   celprp (jp_build_list), dnergy (grad zeroed, then jp_bonded and
!   jp_nonbond accumulate; FP64 and FP32/fixed-point pair arithmetic), and the
!   MD loop (jp_run_nve; METHOD= MICR, or METHOD= CANO at SETTEM = settem K with
!   TEMPCO= NO through jp_set_canonical), and writes every result to an output file.
!
! usage: jp_host <system file> <output file> <nsteps> <update> <nmon> <precision 0|1> [settem]
!        jp_host --mdinp <md.inp> [system file] [precision 0|1]
!                 md_run's own control file.  EXE> INPUT: TOPOLO (TPL),
!                 COORDI with CRDFMT= PDB (default) or BINA, SETSHK, and CLUSTA= a<b (atoms
!                 a..b in cluster 1, the rest in cluster 2) with GEPSME= LMD,
!                 read by jp_input.f90.  EXE> MD: LOOPLI, UPDATE, TIMEST,
!                 METHOD (MICR, CANO or GE; required), SETTEM, TEMPCO, HEATLO, STARTT, OUTLOG,
!                 STOPCE, NSTPCN (jp_set_stopce, the chains of the TPL),
!                 RESTAR (NO: md_run's initial half step, jp_start_step, from
!                 INITIA= ZERO), NAMERI, NAMERO, NAMETO, NAMEGE, LAMBDA, LMDSQU, CUTLEN,
!                 DIEVAL, XMINCL..ZMAXCL, COVSHK, LIMSHK (inpmd.f90; defaults of
!                 main.f90:150-185).  CAL15M= ZD runs under BOUNDA= PBC;
!                 CAL15M= CMM (md_run's default) under BOUNDA= NO (its default)
!                 with EXE> INPUT's CELSIZ, CELLVL and EXE> MD's CMMUPD
!                 (jp_set_cmm).  METHOD= GE reads the NAMEGE= parameter
!                 file (inmcmd.f90): multicanonical MD (jp_set_mcmd), or under
!                 GEPSME= LMD adaptive lambda square dynamics (jp_set_alsd, with
!                 EXE> INPUT's LWEIGH and LTCNTL).
!                 NAMETO receives the potential energy of every step per
!                 energy class in md_run's format (md.f90:578-601).
!                 x, v and lambda from the NAMERI restart, the final state
!                 written to NAMERO in md_run's restart format
!                 (outrst.f90:56-61).  Given a system file, the topology and
!                 the run parameters other than TIMEST come from it instead
!                 (and TEMPCO= YES is refused: it carries no molecule kinds).
!        jp_host --tplsys <md.inp> <system file>
!                 the system md.inp's TPL/SHK/BINA inputs describe, written as
!                 a system file (x and v from the restart with RESTAR= YES);
!                 no GPU use
!        jp_host --layout   (byte sizes and offsets of the bind(C) types; no GPU use)
!
! System file (Fortran unformatted sequential, one array per record):
!   natom ntyp nfrag nfrg nexcl nexcess nshake shake_maxit cell_level lmd nbond nangle ntors n14 (int32)
!   shake_tol cutoff diel dt lambda celwal(6)                               (real*8)
!   cord(3,N) vel(3,N) mass(N) charge(N) atype(N) icls(N) rmin2(ntyp) eps(ntyp)
!   n_matrix(nfrag,nfrag) excl(2,nexcl) excess(2,nexcess)
!   shake_natom(nshake) shake_atoms(4,nshake) shake_dist2(6,nshake)
!   bond_atoms(2,nbond) bond_fc(nbond) bond_r0(nbond)
!   angle_atoms(3,nangle) angle_fc(nangle) angle_theta0(nangle)
!   tors_atoms(4,ntors) tors_fc tors_rot tors_phase(ntors) tors_div(ntors)
!   p14_atoms(2,n14) p14_vdw(n14) p14_elec(n14)
! Output file (same record format):
!   ier(9): create, build_list, nonbond FP64, nonbond FP32, run_nve, get_state, bonded (x2),
!           set_canonical (with --mdinp, ier(7) is the restart step's and
!           ier(8) METHOD= GE's set-up: jp_set_mcmd or jp_set_alsd)
!   grad64(3,N,nfrg) eelec64(nfrg) evdw64(nfrg) grad32(3,N,nfrg) eelec32(nfrg) evdw32(nfrg)
!   eterm(6,nfrg) = enevcc(2:7,:) (bond, angle, torsion, improper, 1-4 vdW, 1-4 elec)
!   cord(3,N) vel(3,N) after the run, nlog, then per record: step etot epot ekin
program jp_host
  use iso_c_binding
  use jpgpu
  use jp_input
  implicit none
  character(len=1024) :: fin, fout, arg
  integer(c_int32_t) :: hdr(14), nsteps, update, nmon, prec
  real(c_double) :: rhdr(11)
  integer :: n, ntyp, nfrag, nfrg, nexcl, nexcess, nshake, nbond, nangle, ntors, n14, nlog, k, u
  integer(c_int) :: ier(9)
  real(c_double) :: settem, timest
  character(len=1024) :: mdinp, rstin, rstout, etof
  logical :: mdmode, restar, tempco, tplsys, fromtpl
  character(len=1024) :: tplf, crdf, shkf, crdfmt
  integer :: cla, clb
  logical :: lmdin
  real(c_double) :: cutlen, dieval, box(6), covshk
  integer(c_int32_t) :: limshk, cellvl
  real(c_double) :: cputim
  logical :: initzero
  ! CAL15M= CMM (md_run's default) with BOUNDA= NO: CELSIZ, CELLVL, CMMUPD
  logical :: cmm
  real(c_double) :: celsiz
  integer(c_int32_t) :: cmmupd
  ! the heating loop of METHOD= CANO / GE (HEATLO steps from STARTT K)
  integer(c_int32_t) :: heatlo
  real(c_double) :: startt
  ! STOPCE= TRAN/ROTA/BOTH (1/2/3; 0 NO) of the first NSTPCN chains
  integer(c_int32_t) :: stopce, nstpcn
  integer(8) :: tstart, trate, tnow          ! wall clock since the program started (md_run's fxcpus)
  integer(c_int32_t), allocatable, target :: grp(:)
  type(top_t) :: top
  character(len=80) :: title
  integer(c_int32_t) :: rn, rnvar, rloop
  real(c_double) :: rtime, ret, rek, rep, rlam, rlamv
  ! METHOD= GE: the NAMEGE= file (inmcmd.f90:20-41); LAMBDA/LMDSQU; LWEIGH, LTCNTL
  logical :: gemode
  character(len=1024) :: gef
  integer :: ge_nw, ge_nd
  real(c_double), allocatable, target :: ge_c(:,:), ge_low(:), ge_high(:)
  real(c_double) :: ge_alw, ge_aup, ge_celw, ge_ceup, ge_tempce, ge_lower, ge_upper
  real(c_double) :: lweigh, lamkw, lsqkw, lamv0
  integer(c_int32_t) :: ltcntl
  real(c_double), allocatable, target :: cord(:,:), vel(:,:), mass(:), charge(:), rmin2(:), eps(:)
  real(c_double), allocatable, target :: sdist(:,:), grad64(:,:,:), grad32(:,:,:)
  real(c_double), target :: eel64(3), evd64(3), ehy64(3), eel32(3), evd32(3), ehy32(3), eterm(6,3)
  real(c_double), allocatable, target :: bfc(:), br0(:), afc(:), ath(:), tfc(:), trot(:), tph(:), pv(:), pe(:)
  integer(c_int32_t), allocatable, target :: batm(:,:), aatm(:,:), tatm(:,:), tdiv(:), patm(:,:)
  integer(c_int32_t), allocatable, target :: atype(:), icls(:), nmat(:,:), excl(:,:), excess(:,:)
  integer(c_int32_t), allocatable, target :: snat(:), satm(:,:)
  type(jp_energy_log), allocatable, target :: elog(:)
  type(jp_sys_desc), target :: desc
  type(jp_energy_log), target :: rec1
  type(c_ptr) :: ctx

  call get_command_argument(1, arg)
  if (trim(arg) == '--layout') then
     write(*,'(a,i0)') 'sizeof jp_sys_desc ', c_sizeof(desc)
     write(*,'(a,i0)') 'offset excl ', off(c_loc(desc%excl))
     write(*,'(a,i0)') 'offset shake_tol ', off(c_loc(desc%shake_tol))
     write(*,'(a,i0)') 'offset celwal ', off(c_loc(desc%celwal))
     write(*,'(a,i0)') 'offset lmd ', off(c_loc(desc%lmd))
     write(*,'(a,i0)') 'offset lambda ', off(c_loc(desc%lambda))
     write(*,'(a,i0)') 'offset nbond ', off(c_loc(desc%nbond))
     write(*,'(a,i0)') 'offset angle_cls ', off(c_loc(desc%angle_cls))
     write(*,'(a,i0)') 'sizeof jp_energy_log ', c_sizeof(rec1)
     stop
  end if
  tplsys = trim(arg) == '--tplsys'
  call system_clock(tstart, trate)
  mdmode = trim(arg) == '--mdinp' .or. tplsys
  settem = 0.0d0
  fin = ' '; fout = ' '
  if (mdmode) then
     call get_command_argument(2, mdinp)
     prec = 0
     do k = 3, command_argument_count()
        call get_command_argument(k, arg)
        if (tplsys) then
           if (k == 3) fout = arg
        else if (trim(arg) == '0' .or. trim(arg) == '1') then
           read(arg, *) prec
        else
           fin = arg
        end if
     end do
     if (len_trim(mdinp) == 0 .or. (tplsys .and. len_trim(fout) == 0)) then
        write(*,'(a)') 'usage: jp_host --mdinp <md.inp> [system file] [precision 0|1]'
        write(*,'(a)') '       jp_host --tplsys <md.inp> <system file>'
        stop 2
     end if
     call read_mdinp(trim(mdinp))
  else
     if (command_argument_count() < 6) then
        write(*,'(a)') 'usage: jp_host <system file> <output file> <nsteps> <update> <nmon> <precision 0|1>'
        stop 2
     end if
     call get_command_argument(1, fin)
     call get_command_argument(2, fout)
     call get_command_argument(3, arg); read(arg, *) nsteps
     call get_command_argument(4, arg); read(arg, *) update
     call get_command_argument(5, arg); read(arg, *) nmon
     call get_command_argument(6, arg); read(arg, *) prec
     if (command_argument_count() >= 7) then
        call get_command_argument(7, arg); read(arg, *) settem
     end if
  end if

  fromtpl = len_trim(fin) == 0
  if (fromtpl) then
     call load_inputs()
  else
     call read_system_file()
  end if

  desc%natom = n; desc%ntyp = ntyp; desc%nfrag = nfrag; desc%nfrg = nfrg
  desc%x = c_loc(cord); desc%mass = c_loc(mass); desc%charge = c_loc(charge)
  desc%atype = c_loc(atype); desc%rmin2 = c_loc(rmin2); desc%eps = c_loc(eps)
  desc%icls = c_loc(icls); desc%n_matrix = c_loc(nmat)
  desc%nexcl = nexcl; desc%excl = c_loc(excl)
  desc%nexcess = nexcess; desc%excess = c_loc(excess)
  desc%nshake = nshake; desc%shake_natom = c_loc(snat)
  desc%shake_atoms = c_loc(satm); desc%shake_dist2 = c_loc(sdist)
  desc%shake_maxit = hdr(8); desc%cell_level = hdr(9); desc%lmd = hdr(10)
  desc%shake_tol = rhdr(1); desc%cutoff = rhdr(2); desc%diel = rhdr(3); desc%dt = rhdr(4)
  desc%lambda = rhdr(5); desc%celwal = rhdr(6:11)
  desc%nbond = nbond; desc%bond_atoms = c_loc(batm); desc%bond_fc = c_loc(bfc); desc%bond_r0 = c_loc(br0)
  desc%nangle = nangle; desc%angle_atoms = c_loc(aatm); desc%angle_fc = c_loc(afc)
  desc%angle_theta0 = c_loc(ath)
  desc%ntors = ntors; desc%tors_atoms = c_loc(tatm); desc%tors_fc = c_loc(tfc); desc%tors_rot = c_loc(trot)
  desc%tors_phase = c_loc(tph); desc%tors_div = c_loc(tdiv)
  desc%n14 = n14; desc%p14_atoms = c_loc(patm); desc%p14_vdw = c_loc(pv); desc%p14_elec = c_loc(pe)

  if (mdmode) then
     desc%dt = timest
     ! lambda (inpmd.f90:290-297: LAMBDA, else sqrt(LMDSQU); inprst.f90:56-66:
     ! the restart's unless set, its velocity only under METHOD= GE)
     if (lmdin .and. lamkw /= 0.0d0) then
        desc%lambda = lamkw
     else if (lmdin .and. lsqkw /= 0.0d0) then
        desc%lambda = sqrt(lsqkw)
     end if
     lamv0 = 0.0d0
     if (restar) then
        call read_restart(trim(rstin))
        if (fromtpl .and. desc%lambda == 0.0d0) desc%lambda = rlam
        if (gemode) lamv0 = rlamv
     end if
     if (cmm) then
        ! CMM has no box: the context's walls only enclose the atoms (their
        ! cells follow the coordinates, celpar_CB)
        do k = 1, 3
           desc%celwal(2 * k - 1) = minval(cord(k, :)) - cutlen - 1.0d0
           desc%celwal(2 * k) = max(maxval(cord(k, :)) + cutlen + 1.0d0, desc%celwal(2 * k - 1) + 2.0d0 * cutlen + 2.0d0)
        end do
     end if
     if (tplsys) then
        call write_system_file()
        stop
     end if
  end if

  write(*,'(2a)') 'jp_host: ', jp_version()
  ier = 0
  ier(1) = jp_ctx_create(desc, ctx)
  if (ier(1) /= JP_OK) then
     write(*,'(a,i0)') 'jp_ctx_create: ier = ', ier(1)
     stop 1
  end if
  ier(1) = jp_set_state(ctx, c_null_ptr, c_loc(vel))
  if (mdmode .and. cmm) then
     ier(1) = jp_set_cmm(ctx, 1_c_int32_t, celsiz, cellvl, cmmupd)   ! incel's CMM set-up
     if (ier(1) /= JP_OK) then
        write(*,'(a,i0)') 'jp_set_cmm: ier = ', ier(1)
        stop 1
     end if
  end if
  ier(2) = jp_build_list(ctx, c_null_ptr)                 ! celprp
  if (mdmode) then
     call run_mdinp()
     stop
  end if

  ! dnergy: grad zeroed, then bonds and angles, then cmmdo (dnergy.f90:45-55, dnergv.f90)
  allocate(grad64(3,n,nfrg), grad32(3,n,nfrg))
  grad64 = 0.0d0; eel64 = 0.0d0; evd64 = 0.0d0; ehy64 = 0.0d0; eterm = 0.0d0
  ier(7) = jp_bonded(ctx, c_null_ptr, c_loc(grad64), c_loc(eterm))
  ier(3) = jp_nonbond(ctx, c_null_ptr, c_loc(grad64), c_loc(eel64), c_loc(evd64), c_loc(ehy64), JP_PREC_FP64)
  grad32 = 0.0d0; eel32 = 0.0d0; evd32 = 0.0d0; ehy32 = 0.0d0
  ier(8) = jp_bonded(ctx, c_null_ptr, c_loc(grad32), c_null_ptr)
  ier(4) = jp_nonbond(ctx, c_null_ptr, c_loc(grad32), c_loc(eel32), c_loc(evd32), c_loc(ehy32), JP_PREC_FP32_FIXED)

  ! the MD loop: METHOD= MICR, or METHOD= CANO (md.f90:546-552, verlethe)
  if (settem > 0.0d0) ier(9) = jp_set_canonical(ctx, 1_c_int32_t, settem, 3_c_int32_t, 1.0d-3, c_null_ptr)
  nlog = 0
  if (nmon > 0) nlog = nsteps / nmon
  allocate(elog(nlog + 1))
  ier(5) = jp_run_nve(ctx, nsteps, update, nmon, prec, c_loc(elog))
  ier(6) = jp_get_state(ctx, c_loc(cord), c_loc(vel))
  write(*,'(a,9(1x,i0))') 'jp_host: ier =', ier
  if (nlog > 0) write(*,'(a,i0,a,f16.6)') 'jp_host: step ', elog(nlog)%step, ' etot ', elog(nlog)%etot

  open(newunit=u, file=trim(fout), form='unformatted', access='sequential', status='replace', action='write')
  write(u) ier
  write(u) grad64
  write(u) eel64(1:nfrg)
  write(u) evd64(1:nfrg)
  write(u) grad32
  write(u) eel32(1:nfrg)
  write(u) evd32(1:nfrg)
  write(u) eterm(:, 1:nfrg)
  write(u) cord
  write(u) vel
  write(u) nlog
  do k = 1, nlog
     write(u) elog(k)%step, elog(k)%etot, elog(k)%epot, elog(k)%ekin
  end do
  close(u)
  ier(1) = jp_ctx_destroy(ctx)

contains

  ! the system file (record layout above)
  subroutine read_system_file()
    open(newunit=u, file=trim(fin), form='unformatted', access='sequential', status='old', action='read')
    read(u) hdr
    read(u) rhdr
    n = hdr(1); ntyp = hdr(2); nfrag = hdr(3); nfrg = hdr(4)
    nexcl = hdr(5); nexcess = hdr(6); nshake = hdr(7); nbond = hdr(11); nangle = hdr(12)
    ntors = hdr(13); n14 = hdr(14)
      call alloc_arrays()
    read(u) cord
    read(u) vel
    read(u) mass
    read(u) charge
    read(u) atype
    read(u) icls
    read(u) rmin2
    read(u) eps
    read(u) nmat
    read(u) excl(:, 1:nexcl)
    read(u) excess(:, 1:nexcess)
    read(u) snat(1:nshake)
    read(u) satm(:, 1:nshake)
    read(u) sdist(:, 1:nshake)
    read(u) batm(:, 1:nbond)
    read(u) bfc(1:nbond)
    read(u) br0(1:nbond)
    read(u) aatm(:, 1:nangle)
    read(u) afc(1:nangle)
    read(u) ath(1:nangle)
    read(u) tatm(:, 1:ntors)
    read(u) tfc(1:ntors)
    read(u) trot(1:ntors)
    read(u) tph(1:ntors)
    read(u) tdiv(1:ntors)
    read(u) patm(:, 1:n14)
    read(u) pv(1:n14)
    read(u) pe(1:n14)
    close(u)
  end subroutine read_system_file

  subroutine alloc_arrays()
    allocate(cord(3,n), vel(3,n), mass(n), charge(n), atype(n), icls(n), rmin2(ntyp), eps(ntyp))
    allocate(nmat(nfrag,nfrag), excl(2,max(nexcl,1)), excess(2,max(nexcess,1)))
    allocate(snat(max(nshake,1)), satm(4,max(nshake,1)), sdist(6,max(nshake,1)))
    allocate(batm(2,max(nbond,1)), bfc(max(nbond,1)), br0(max(nbond,1)))
    allocate(aatm(3,max(nangle,1)), afc(max(nangle,1)), ath(max(nangle,1)))
    allocate(tatm(4,max(ntors,1)), tfc(max(ntors,1)), trot(max(ntors,1)), tph(max(ntors,1)), tdiv(max(ntors,1)))
    allocate(patm(2,max(n14,1)), pv(max(n14,1)), pe(max(n14,1)))
  end subroutine alloc_arrays

  ! EXE> INPUT and EXE> MD sections of md_run's control file (inpmd.f90
  ! keywords, main.f90:150-185 defaults; METHOD must be given)
  subroutine read_mdinp(path)
    character(len=*), intent(in) :: path
    character(len=256) :: line, rest, key, val
    integer :: uu, ios, p, q
    logical :: inmd, inin, have_method, have_pbc
    integer :: iv
    real(c_double) :: rv
    nsteps = 0; update = 10; nmon = 0; timest = 2.0d0; settem = -300.0d0   ! negative: METHOD= MICR
    restar = .false.; tempco = .true.; rstin = ' '; rstout = ' '; etof = ' '
    tplf = ' '; crdf = ' '; shkf = ' '; crdfmt = 'PDB'; cla = 0; clb = 0; lmdin = .false.
    cutlen = 12.0d0; dieval = 1.0d0; box = 0.0d0; covshk = 1.0d-6; limshk = 1000
    cellvl = 3; cputim = 99999999.0d0; initzero = .false.
    cmm = .true.; celsiz = 6.0d0; cmmupd = 5   ! main.f90:170, inpinp.f90:86, inpmd.f90:151
    heatlo = 0; startt = 300.0d0                 ! inpmd.f90:127,147
    stopce = 0; nstpcn = 1                       ! inpmd.f90:124, intele(84)
    gemode = .false.; gef = ' '; lweigh = 0.0d0; ltcntl = 1; lamkw = 0.0d0; lsqkw = 0.0d0
    have_method = .false.; have_pbc = .false.
    inmd = .false.; inin = .false.
    open(newunit=uu, file=path, status='old', action='read')
    do
       read(uu, '(a)', iostat=ios) line
       if (ios /= 0) exit
       if (index(line, 'EXE>') > 0) then
          inmd = index(line, 'EXE> MD') > 0
          inin = index(line, 'EXE> INPUT') > 0
          cycle
       end if
       if (.not. (inmd .or. inin)) cycle
       if (index(line, 'QUIT') > 0) then
          inmd = .false.; inin = .false.
          cycle
       end if
       rest = line
       do
          p = index(rest, '=')
          if (p < 2) exit
          key = adjustl(rest(max(1, p - 6):p - 1))
          rest = adjustl(rest(p + 1:))
          q = index(rest, ' ')
          if (q == 0) q = len_trim(rest) + 1
          val = rest(1:q - 1)
          rest = rest(q:)
          ios = 0
          if (inin) then
             ! EXE> INPUT keywords (inpinp.f90:40-44)
             select case (trim(key))
             case ('TOPOLO'); tplf = val
             case ('COORDI'); crdf = val
             case ('CRDFMT'); crdfmt = val
             case ('SETSHK'); shkf = val
             case ('TPLFMT')
                if (val(1:4) /= 'FORM') call refuse(key, val, 'only the formatted TPL is read')
             case ('GEPSME')
                if (trim(val) == 'LMD') then
                   lmdin = .true.
                else if (trim(val) /= 'AND') then
                   call refuse(key, val, 'cluster methods other than AND and LMD are not implemented')
                end if
             case ('CELLVL')
                read(val, *, iostat=ios) cellvl
                if (ios == 0 .and. (cellvl < 1 .or. cellvl > 6)) call refuse(key, val, 'cell levels 1-6 only')
             case ('CELSIZ')   ! the level-1 cell of CMM (ZD sets its own, incel.f90:29-34)
                read(val, *, iostat=ios) celsiz
                if (ios == 0 .and. .not. (celsiz > 0.0d0)) call refuse(key, val, 'a positive cell size')
             case ('SCALET')
                if (trim(val) /= '0030') call refuse(key, val, 'only the default class choice 0030 is implemented')
             case ('LWEIGH')   ! lambda's mass (0: the class-1 atoms' mass), ALSD only
                read(val, *, iostat=ios) lweigh
             case ('LTCNTL')   ! lambda's temperature control, ALSD only (inpinp.f90:178-182)
                if (trim(val) == 'NO') then
                   ltcntl = 0
                else if (trim(val) == 'YES') then
                   ltcntl = 1
                else
                   call refuse(key, val, 'NO or YES')
                end if
             case ('REFCOO', 'OUTMON')
                write(*,'(3a)') 'jp_host: ', trim(key), ' (monitoring only) is not written by this host'
             case ('SETVAR', 'SETBOU', 'POSITI', 'DISTAN', 'DIHEDR', 'SETORI', 'EXTCMM', 'REPULS', &
                   'CELRES', 'CLUSTB')
                call refuse(key, val, 'not implemented by this host')
             case default
                call refuse(key, val, 'not an EXE> INPUT keyword of md_run')
             case ('CLUSTA')
                q = index(val, '<')
                if (q < 2) then
                   write(*,'(2a)') 'jp_host: CLUSTA must read a<b: ', trim(val)
                   stop 3
                end if
                read(val(1:q - 1), *, iostat=ios) cla
                if (ios == 0) read(val(q + 1:), *, iostat=ios) clb
             end select
             if (ios /= 0) call bad_value(key, val)
             cycle
          end if
          ! EXE> MD keywords (inpmd.f90:37-50): implemented, accepted as
          ! no-ops, output-only (warned), or refused when their value would make
          ! the run differ from md_run's
          select case (trim(key))
          case ('CAL15M')
             if (trim(val) == 'ZD') then
                cmm = .false.
             else if (trim(val) /= 'CMM') then
                call refuse(key, val, 'only zero-dipole (ZD) and the cell multipole method (CMM) are implemented')
             end if
          case ('BOUNDA')
             if (trim(val) == 'PBC') then
                have_pbc = .true.
             else if (trim(val) /= 'NO') then
                call refuse(key, val, 'only periodic boundaries (PBC, with ZD) and none (NO, with CMM) are implemented')
             end if
          case ('CMMUPD'); read(val, *, iostat=ios) cmmupd
          case ('STOPCE')
             select case (trim(val))
             case ('NO'); stopce = 0
             case ('TRAN'); stopce = 1
             case ('ROTA'); stopce = 2
             case ('BOTH'); stopce = 3
             case default; call refuse(key, val, 'TRAN, ROTA, BOTH or NO')
             end select
          case ('NSTPCN')
             read(val, *, iostat=ios) nstpcn
             if (ios == 0 .and. nstpcn < 1) call refuse(key, val, 'a chain count >= 1')
          case ('HEATLO')
             read(val, *, iostat=ios) heatlo
             if (ios == 0 .and. heatlo < 0) call refuse(key, val, 'a step count >= 0')
          case ('STARTT'); read(val, *, iostat=ios) startt
          case ('SHAKEM')
             if (trim(val) /= 'HBON') call refuse(key, val, 'only SHAKEM= HBON (the SHK file''s groups) is implemented')
          case ('PSETMP')
             read(val, *, iostat=ios) rv
             if (ios == 0 .and. rv > 1.0d-10) call refuse(key, val, 'pseudo-temperature scaling is not implemented')
          case ('LAMBDA'); read(val, *, iostat=ios) lamkw   ! inpmd.f90:290-297
          case ('LMDSQU'); read(val, *, iostat=ios) lsqkw
          case ('NAMEGE'); gef = val
          case ('LXCELL', 'LYCELL', 'LZCELL', 'SETCEN', 'CENTRX', 'CENTRY', 'CENTRZ')
             call refuse(key, val, 'give the box as XMINCL..ZMAXCL')
          case ('CPUTIM'); read(val, *, iostat=ios) cputim
          case ('INITIA')
             if (trim(val) /= 'ZERO' .and. trim(val) /= 'SET') call refuse(key, val, 'ZERO or SET')
             initzero = trim(val) == 'ZERO'
          case ('MNTRTO')
             if (trim(val) == 'NO') then
                etof = ' '
             else if (trim(val) /= 'ASCI') then
                call refuse(key, val, 'NAMETO is written as ASCII only')
             end if
          case ('PARALV', 'N_SIMD', 'LOGFOR', 'RELAXA', 'TEMPER', 'OUTATM', 'DUMMYL', 'BESTFI', &
                'RANDOM', 'RADCAP', 'FORCAP', 'FUNCAP', 'WETPSR', 'WETDSR', 'WETDHR', &
                'CAPSHP', 'FOACAP', 'RADPRO', 'HTPSLO', 'STAPST', 'ELLIPA', 'ELLIPB', 'ELLIPC', 'ELLBUF', &
                'CELBUF', 'RADIUS')
             ! thread layout and SIMD width, log format, methods and boundaries
             ! this host does not use (CONS relaxation, random initial
             ! velocities, the CAP and ellipsoid parameters, which act only
             ! through SETBOU=, refused above): no effect on the run
          case ('NAMECO', 'NAMEVE', 'NAMEEN', 'NAMETR', 'NAMECL', 'OUTCOO', 'OUTVEL', 'OUTENE', &
                'OUTTRJ', 'OUTCLE', 'MNTRCO', 'MNTRVE', 'MNTREN', 'MNTRTR', 'MNTRCL', 'FDEBUG')
             write(*,'(3a)') 'jp_host: ', trim(key), ' (trajectory/monitoring output) is not written by this host'
          case ('LOOPLI'); read(val, *, iostat=ios) nsteps
          case ('UPDATE'); read(val, *, iostat=ios) update
          case ('TIMEST'); read(val, *, iostat=ios) timest
          case ('OUTLOG'); read(val, *, iostat=ios) nmon
          case ('SETTEM'); read(val, *, iostat=ios) rep; settem = sign(rep, settem)
          case ('RESTAR'); restar = val(1:3) == 'YES'
          case ('NAMERI'); rstin = val
          case ('NAMERO'); rstout = val
          case ('NAMETO'); etof = val
          case ('TEMPCO'); tempco = val(1:3) == 'YES'
          case ('CUTLEN'); read(val, *, iostat=ios) cutlen
          case ('DIEVAL'); read(val, *, iostat=ios) dieval
          case ('COVSHK'); read(val, *, iostat=ios) covshk
          case ('LIMSHK'); read(val, *, iostat=ios) limshk
          case ('XMINCL'); read(val, *, iostat=ios) box(1)
          case ('XMAXCL'); read(val, *, iostat=ios) box(2)
          case ('YMINCL'); read(val, *, iostat=ios) box(3)
          case ('YMAXCL'); read(val, *, iostat=ios) box(4)
          case ('ZMINCL'); read(val, *, iostat=ios) box(5)
          case ('ZMAXCL'); read(val, *, iostat=ios) box(6)
          case ('METHOD')
             have_method = .true.
             if (val(1:4) == 'CANO') then
                settem = abs(settem)
             else if (val(1:2) == 'GE') then
                settem = abs(settem)
                gemode = .true.
             else if (val(1:4) == 'MICR') then
                settem = -abs(settem)
             else
                call refuse(key, val, 'only METHOD= MICR, CANO and GE are implemented')
             end if
          case default
             call refuse(key, val, 'not an EXE> MD keyword of md_run')
          end select
          if (ios /= 0) call bad_value(key, val)
       end do
    end do
    close(uu)
    if (.not. have_method) then
       write(*,'(a)') 'jp_host: METHOD= MICR, CANO or GE must be given in EXE> MD'
       stop 3
    end if
    if (gemode) then
       if (len_trim(gef) == 0) then
          write(*,'(a)') 'jp_host: METHOD= GE needs its parameter file (NAMEGE=)'
          stop 3
       end if
       call read_namege(trim(gef))
    end if
    ! md_run's defaults are CMM electrostatics and no boundary (main.f90:149,170);
    ! ZD runs here under periodic boundaries, CMM without
    if (.not. cmm .and. .not. have_pbc) then
       write(*,'(a)') 'jp_host: CAL15M= ZD needs BOUNDA= PBC (md_run''s default is no boundary)'
       stop 3
    end if
    if (cmm .and. have_pbc) then
       write(*,'(a)') 'jp_host: CAL15M= CMM is implemented for BOUNDA= NO only'
       stop 3
    end if
    if (cmm .and. (cmmupd < 1 .or. mod(update, cmmupd) /= 0)) then
       write(*,'(a)') 'jp_host: ERROR> INPMD: MOD(UPDATE,CMMUPD) should be 0'   ! inpmd.f90:284-287
       stop 3
    end if
    if (.not. restar .and. .not. initzero .and. len_trim(fin) == 0) then
       write(*,'(a)') 'jp_host: RESTAR= NO needs INITIA= ZERO (md_run''s random initial velocities are not reproduced)'
       stop 3
    end if
    if (nmon <= 0 .or. nmon > nsteps) nmon = 0
    if (stopce > 0 .and. len_trim(fin) > 0) then
       write(*,'(a)') 'jp_host: STOPCE needs the chains of the TPL (a system file has none)'
       stop 3
    end if
    if (settem > 0.0d0 .and. tempco .and. len_trim(fin) > 0) then
       write(*,'(a)') 'jp_host: TEMPCO= YES needs the molecule kinds of the TPL (a system file has none)'
       stop 3
    end if
  end subroutine read_mdinp

  ! inmcmd.f90:20-41: nwindow ndeg; per window ndeg+1 coefficients, one per
  ! record, then low_v high_v; alphalw alphaup; celw ceup; tempce; under
  ! GEPSME= LMD also lower upper.  List-directed reads, as md_run's.
  subroutine read_namege(path)
    character(len=*), intent(in) :: path
    integer :: ug, ios, i, j
    open(newunit=ug, file=path, status='old', action='read', iostat=ios)
    if (ios /= 0) then
       write(*,'(2a)') 'jp_host: cannot open the NAMEGE= file ', path
       stop 3
    end if
    read(ug, *, iostat=ios) ge_nw, ge_nd
    if (ios /= 0 .or. ge_nw < 1 .or. ge_nd < 0) call bad_ge()
    allocate(ge_c(0:ge_nd, ge_nw), ge_low(ge_nw), ge_high(ge_nw))
    do i = 1, ge_nw
       do j = 0, ge_nd
          read(ug, *, iostat=ios) ge_c(j, i)
          if (ios /= 0) call bad_ge()
       end do
       read(ug, *, iostat=ios) ge_low(i), ge_high(i)
       if (ios /= 0) call bad_ge()
    end do
    read(ug, *, iostat=ios) ge_alw, ge_aup
    if (ios == 0) read(ug, *, iostat=ios) ge_celw, ge_ceup
    if (ios == 0) read(ug, *, iostat=ios) ge_tempce
    if (ios /= 0) call bad_ge()
    ge_lower = 0.0d0; ge_upper = 0.0d0
    if (lmdin) then
       read(ug, *, iostat=ios) ge_lower, ge_upper
       if (ios /= 0) call bad_ge()
    end if
    close(ug)
  contains
    subroutine bad_ge()
      write(*,'(2a)') 'jp_host: READ FILE ERROR IN the NAMEGE= file ', path
      stop 3
    end subroutine bad_ge
  end subroutine read_namege

  subroutine refuse(key, val, why)
    character(len=*), intent(in) :: key, val, why
    write(*,'(6a)') 'jp_host: ', trim(key), '= ', trim(val), ': ', why
    stop 3
  end subroutine refuse

  subroutine bad_value(key, val)
    character(len=*), intent(in) :: key, val
    write(*,'(4a)') 'jp_host: cannot read the value of ', trim(key), ': ', trim(val)
    stop 3
  end subroutine bad_value

  ! md_run restart file (outrst.f90:56-61): title, natom nvar, loop time
  ! etot ekin epot, cord(3,natom), vel(3,nvar), lambda lambda_v
  subroutine read_restart(path)
    character(len=*), intent(in) :: path
    integer :: uu
    open(newunit=uu, file=path, form='unformatted', access='sequential', status='old', action='read')
    read(uu) title
    read(uu) rn, rnvar
    if (rn /= n .or. rnvar /= n) then
       write(*,'(a,2(1x,i0))') 'jp_host: restart atoms / moving atoms differ from the system:', rn, rnvar
       stop 3
    end if
    read(uu) rloop, rtime, ret, rek, rep
    read(uu) cord
    read(uu) vel
    read(uu) rlam, rlamv
    close(uu)
  end subroutine read_restart

  subroutine run_mdinp()
    integer :: uu, nl, nm, kk, ndone, nseg
    real(c_double) :: epc(3)
    if (.not. restar) then
       rloop = 0; rtime = 0.0d0
    end if
    if (stopce > 0) then
       ! remotv over ixcend(NSTPCN) atoms; the degrees of freedom of atom 1's
       ! molecule kind (md.f90:114-121, 183-215) before the canonical set-up
       if (nstpcn > size(top%cend) .or. top%cend(min(nstpcn, size(top%cend))) <= 0) then
          write(*,'(a,i0)') 'jp_host: NSTPCN beyond the chains of the TPL: ', nstpcn
          stop 3
       end if
       kk = 1
       do while (kk < n)
          if (top%kind(kk + 1) /= top%kind(1)) exit
          kk = kk + 1
       end do
       if (jp_set_stopce(ctx, stopce, top%cend(nstpcn), int(kk, c_int32_t)) /= JP_OK) then
          write(*,'(a)') 'jp_host: jp_set_stopce failed'
          stop 1
       end if
    end if
    if (settem > 0.0d0) then
       if (tempco .and. fromtpl) then
          call temperature_groups()
          ier(9) = jp_set_canonical(ctx, 1_c_int32_t, settem, 3_c_int32_t, 1.0d-3, c_loc(grp))
       else
          ier(9) = jp_set_canonical(ctx, 1_c_int32_t, settem, 3_c_int32_t, 1.0d-3, c_null_ptr)
       end if
       ! the heating loop (verlethe.f90:419-423): the target ramps from STARTT
       if (ier(9) == 0 .and. heatlo > 0) ier(9) = jp_set_heating(ctx, heatlo, startt)
    end if
    if (gemode .and. ier(9) == 0) then
       if (lmdin) then   ! ALSD (verlethe.f90:235-290)
          ier(8) = jp_set_alsd(ctx, 1_c_int32_t, int(ge_nw, c_int32_t), int(ge_nd, c_int32_t), c_loc(ge_c), &
                               c_loc(ge_low), c_loc(ge_high), ge_lower, ge_upper, ge_celw, ge_ceup, ge_tempce, &
                               lweigh, ltcntl)
          if (ier(8) == 0) ier(8) = jp_set_lambda(ctx, desc%lambda, lamv0)
       else              ! multicanonical MD (verlethe.f90:196-233)
          ier(8) = jp_set_mcmd(ctx, 1_c_int32_t, int(ge_nw, c_int32_t), int(ge_nd, c_int32_t), c_loc(ge_c), &
                               c_loc(ge_low), c_loc(ge_high), ge_alw, ge_aup, ge_celw, ge_ceup, ge_tempce)
       end if
       if (ier(8) /= 0) then
          write(*,'(a,i0)') 'jp_host: the METHOD= GE set-up failed, ier = ', ier(8)
          stop 1
       end if
    end if
    ! NAMETO needs the energies of every step
    nm = nmon
    if (len_trim(etof) > 0) nm = 1
    ! md_run's idloop continues from the restart's loop number (inirst, md.f90:1132)
    if (jp_set_step(ctx, int(rloop, c_int64_t)) /= JP_OK) then
       write(*,'(a,i0)') 'jp_host: jp_set_step failed for loop ', rloop
       stop 1
    end if
    ! a restart starts with inirst's step (md.f90:1104-1263), which keeps the
    ! coordinates and leaves the velocities its last verlethe substep started
    ! from (the restart's own unless a SHAKE retry split the step)
    if (restar) ier(7) = jp_restart_step(ctx, prec, c_null_ptr)
    ! without a restart md_run starts with staver's half step (inista,
    ! md.f90:1056-1065) from the initial velocities
    if (.not. restar) ier(7) = jp_start_step(ctx, prec)
    ndone = nsteps
    if (cputim < 99999999.0d0) then
       ! CPUTIM: md_run checks the elapsed time after every list update and,
       ! past the limit, ends the loop there and writes its restart file
       ! (md.f90:655-676, finmd): run from list update to list update
       allocate(elog(max(nsteps / max(nm, 1) + nsteps / update + 2, 1)))
       nl = 0
       ndone = 0
       do while (ndone < nsteps)
          nseg = min(nsteps - ndone, int(update - mod(int(rloop, 8) + ndone, int(update, 8))))
          ier(5) = jp_run_nve(ctx, nseg, update, nm, prec, c_loc(elog(nl + 1)))
          if (ier(5) /= 0) exit
          ndone = ndone + nseg
          kk = nl
          do while (kk < size(elog))
             if (elog(kk + 1)%step < 0) exit
             kk = kk + 1
          end do
          nl = kk
          call system_clock(tnow)
          if (mod(int(rloop, 8) + ndone, int(update, 8)) == 0 .and. ndone < nsteps .and. &
              real(tnow - tstart, c_double) / real(trate, c_double) >= cputim) then
             write(*,'(a,i0)') 'jp_host: WARNING> MD CPU TIME IS EXCEEDED at loop ', rloop + ndone
             exit
          end if
       end do
    else
       nl = 0
       if (nm > 0) nl = nsteps / nm + 1
       allocate(elog(max(nl, 1)))
       ier(5) = jp_run_nve(ctx, nsteps, update, nm, prec, c_loc(elog))
       ! the records of the monitoring steps come first; the rest have step = -1
       do while (nl > 0)
          if (elog(nl)%step >= 0) exit
          nl = nl - 1
       end do
    end if
    ier(6) = jp_get_state(ctx, c_loc(cord), c_loc(vel))
    write(*,'(a,9(1x,i0))') 'jp_host: ier =', ier
    if (len_trim(etof) > 0 .and. ier(5) == 0) then
       ! md.f90:578-601: epotcc(1,k), the unscaled potential energy of class k
       open(newunit=uu, file=trim(etof), status='replace', action='write')
       do kk = 1, nl
          epc = elog(kk)%ebond + elog(kk)%eangle + elog(kk)%etors + elog(kk)%eimp + elog(kk)%evdw14 &
                + elog(kk)%eelec14 + elog(kk)%evdw + elog(kk)%eelec
          if (desc%lmd /= 0) then
             write(uu, '(2(e15.7,x))', advance='no') elog(kk)%lambda, epc(1)
             if (nfrg == 1) then
                write(uu, *)
             else
                write(uu, '(2(e15.7,x))') epc(2), sum(epc(1:nfrg))
             end if
          else if (nfrg == 1) then
             write(uu, '(e15.7)') epc(1)
          else if (nfrg == 2) then
             write(uu, '(2(e15.7,x))') epc(1), sum(epc(1:2))
          else
             write(uu, '(3(e15.7,x))') epc(1), epc(2), sum(epc(1:3))
          end if
       end do
       close(uu)
    end if
    ret = 0.0d0; rek = 0.0d0; rep = 0.0d0
    if (nl > 0) then
       ret = elog(nl)%etot; rek = elog(nl)%ekin; rep = elog(nl)%epot
    end if
    if (len_trim(rstout) > 0) then
       title = 'jp_host'
       open(newunit=uu, file=trim(rstout), form='unformatted', access='sequential', status='replace', &
            action='write')
       write(uu) title
       write(uu) int(n, c_int32_t), int(n, c_int32_t)
       write(uu) rloop + ndone, rtime + timest * ndone, ret, rek, rep
       write(uu) cord
       write(uu) vel
       ier(8) = jp_get_lambda(ctx, rlam, rlamv)
       if (.not. gemode) rlamv = 0.0d0
       write(uu) rlam, rlamv
       close(uu)
    end if
    if (any(ier /= 0)) stop 1
    ier(1) = jp_ctx_destroy(ctx)
  end subroutine run_mdinp

  ! the system of md.inp's EXE> INPUT files (TPL, SHK, BINA coordinates) and
  ! EXE> MD parameters, in the arrays a system file fills
  subroutine load_inputs()
    integer :: ie
    if (len_trim(tplf) == 0 .or. len_trim(crdf) == 0) then
       write(*,'(a)') 'jp_host: without a system file EXE> INPUT must give TOPOLO and COORDI'
       stop 3
    end if
    if (crdfmt(1:4) /= 'BINA' .and. crdfmt(1:3) /= 'PDB') then
       write(*,'(2a)') 'jp_host: CRDFMT not supported (PDB or BINA): ', trim(crdfmt)
       stop 3
    end if
    if (lmdin .neqv. clb > 0) then
       write(*,'(a)') 'jp_host: GEPSME= LMD and CLUSTA= a<b go together (other cluster methods not supported)'
       stop 3
    end if
    if (.not. cmm .and. (box(2) <= box(1) .or. box(4) <= box(3) .or. box(6) <= box(5))) then
       write(*,'(a)') 'jp_host: EXE> MD must give the periodic box XMINCL..ZMAXCL'
       stop 3
    end if
    call read_topology(trim(tplf), trim(shkf), dieval, top, ie)
    if (ie /= 0) then
       write(*,'(a,i0)') 'jp_host: cannot read TOPOLO / SETSHK, ier = ', ie
       stop 3
    end if
    n = top%n; ntyp = top%ntyp; nexcl = top%nexcl; nexcess = top%nexcess; nshake = top%nshake
    nbond = top%nbond; nangle = top%nangle; ntors = top%ntors; n14 = top%n14
    nfrag = 1; nfrg = 1
    if (lmdin) then
       if (cla < 1 .or. clb > n .or. cla > clb) then
          write(*,'(a,2(1x,i0))') 'jp_host: CLUSTA atoms out of range:', cla, clb
          stop 3
       end if
       nfrag = 2; nfrg = 3
    end if
    call alloc_arrays()
    if (crdfmt(1:4) == 'BINA') then
       call read_crd_bina(trim(crdf), n, cord, ie)
    else
       call read_crd_pdb(trim(crdf), n, cord, ie)
    end if
    if (ie /= 0) then
       write(*,'(a,i0)') 'jp_host: cannot read COORDI (same atom count as the TPL), ier = ', ie
       stop 3
    end if
    vel = 0.0d0
    mass = top%mass; charge = top%chg; atype = top%atype; rmin2 = top%rmin2; eps = top%eps
    icls = 1; nmat = 1
    if (lmdin) then
       ! cluster 1 = atoms cla..clb, cluster 2 = the rest; LMD classes (inpcls.f90:297-324)
       icls = 2; icls(cla:clb) = 1
       nmat = reshape([1, 2, 2, 3], [2, 2])
    end if
    excl(:, 1:nexcl) = top%excl(:, 1:nexcl); excess(:, 1:nexcess) = top%excess(:, 1:nexcess)
    snat(1:nshake) = top%snat(1:nshake); satm(:, 1:nshake) = top%satm(:, 1:nshake)
    sdist(:, 1:nshake) = top%sdist(:, 1:nshake)
    batm(:, 1:nbond) = top%batm(:, 1:nbond); bfc(1:nbond) = top%bfc(1:nbond); br0(1:nbond) = top%br0(1:nbond)
    aatm(:, 1:nangle) = top%aatm(:, 1:nangle); afc(1:nangle) = top%afc(1:nangle)
    ath(1:nangle) = top%ath(1:nangle)
    tatm(:, 1:ntors) = top%tatm(:, 1:ntors); tfc(1:ntors) = top%tfc(1:ntors); trot(1:ntors) = top%trot(1:ntors)
    tph(1:ntors) = top%tph(1:ntors); tdiv(1:ntors) = top%tdiv(1:ntors)
    patm(:, 1:n14) = top%patm(:, 1:n14); pv(1:n14) = top%p14v(1:n14); pe(1:n14) = top%p14e(1:n14)
    hdr = [int(n, c_int32_t), int(ntyp, c_int32_t), int(nfrag, c_int32_t), int(nfrg, c_int32_t), &
           int(nexcl, c_int32_t), int(nexcess, c_int32_t), int(nshake, c_int32_t), limshk, cellvl, &
           merge(1_c_int32_t, 0_c_int32_t, lmdin), int(nbond, c_int32_t), int(nangle, c_int32_t), &
           int(ntors, c_int32_t), int(n14, c_int32_t)]
    rhdr(1) = covshk; rhdr(2) = cutlen; rhdr(3) = dieval; rhdr(4) = timest; rhdr(5) = 0.0d0
    rhdr(6:11) = box
  end subroutine load_inputs

  ! TEMPCO= YES groups (md.f90:136-152): ixamol2 = (molecule kind - 1) *
  ! nfrag + cluster, numbered 1.. without the values no atom takes
  subroutine temperature_groups()
    integer, allocatable :: lab(:)
    integer :: a, g, ng
    allocate(grp(n), lab(maxval(top%kind) * nfrag))
    lab = 0
    do a = 1, n
       lab((top%kind(a) - 1) * nfrag + icls(a)) = 1
    end do
    ng = 0
    do g = 1, size(lab)
       if (lab(g) == 0) cycle
       ng = ng + 1
       lab(g) = ng
    end do
    do a = 1, n
       grp(a) = lab((top%kind(a) - 1) * nfrag + icls(a))
    end do
  end subroutine temperature_groups

  ! the system file of the arrays and parameters in use (x and v included)
  subroutine write_system_file()
    integer :: uw
    rhdr(4) = desc%dt; rhdr(5) = desc%lambda
    open(newunit=uw, file=trim(fout), form='unformatted', access='sequential', status='replace', action='write')
    write(uw) hdr
    write(uw) rhdr
    write(uw) cord
    write(uw) vel
    write(uw) mass
    write(uw) charge
    write(uw) atype
    write(uw) icls
    write(uw) rmin2
    write(uw) eps
    write(uw) nmat
    write(uw) excl(:, 1:nexcl)
    write(uw) excess(:, 1:nexcess)
    write(uw) snat(1:nshake)
    write(uw) satm(:, 1:nshake)
    write(uw) sdist(:, 1:nshake)
    write(uw) batm(:, 1:nbond)
    write(uw) bfc(1:nbond)
    write(uw) br0(1:nbond)
    write(uw) aatm(:, 1:nangle)
    write(uw) afc(1:nangle)
    write(uw) ath(1:nangle)
    write(uw) tatm(:, 1:ntors)
    write(uw) tfc(1:ntors)
    write(uw) trot(1:ntors)
    write(uw) tph(1:ntors)
    write(uw) tdiv(1:ntors)
    write(uw) patm(:, 1:n14)
    write(uw) pv(1:n14)
    write(uw) pe(1:n14)
    close(uw)
  end subroutine write_system_file

  integer function off(p)
    type(c_ptr), intent(in) :: p
    off = int(transfer(p, 0_c_intptr_t) - transfer(c_loc(desc), 0_c_intptr_t))
  end function off

end program jp_host
