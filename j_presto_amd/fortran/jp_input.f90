! Readers of md_run's own input files for the Fortran host (jp_host): the
! TPL topology (inptpl.f90: TITLE, MOLECULES, ATOMS, BONDS, ANGLES, TORSIONS,
! FUNCTIONS, NONBONDS records; ';' comments, '->' continuations), the SHAKE
! file (setshk.f90: SHAKE> SHAKE groups with their distances) and BINA
! (inpcrb.f90) or PDB (inppdb.f90) coordinates.  read_topology expands the molecules by their copy
! counts and derives the tables jp_sys_desc takes, as md_run's input phase
! does: excluded pairs from the ATOMS partner lists (inpcls.f90:229-295), ZD
! excess pairs (incel.f90:665-688), SHAKE groups by size (setshk.f90 modshk),
! the bonds and angles setshk keeps (setshk.f90 reabnd / reaang), the
! torsions incel keeps (incel.f90:993-1009) and the AMBER 1-4 pairs
! (inptpl.f90 mknopm, incel.f90:975-984).
module jp_input
  use iso_c_binding
  implicit none
  private
  public :: read_topology, read_crd_bina, read_crd_pdb, top_t

  integer, parameter :: maxtok = 512, toklen = 64

  type :: mol_t
     character(len=40) :: name = ' '
     integer :: ncopy = 0, na = 0
     real(8), allocatable :: mass(:), chg(:)
     integer, allocatable :: typ(:), npart(:), part(:,:)
     integer :: nb = 0, nang = 0, nt = 0, ns = 0
     integer, allocatable :: b(:,:), a(:,:), t(:,:), tdiv(:), tnbf(:), sn(:), sa(:,:)
     real(8), allocatable :: bfc(:), br0(:), afc(:), ath(:), tfc(:), trot(:), tph(:), sd(:,:)
  end type mol_t

  ! the expanded system, in jp_sys_desc's layout (1-based atom indices)
  type :: top_t
     integer :: n = 0, ntyp = 0, nexcl = 0, nexcess = 0, nshake = 0, nbond = 0, nangle = 0, ntors = 0, n14 = 0
     real(8), allocatable :: mass(:), chg(:), rmin2(:), eps(:), s14e(:), s14v(:)
     integer(c_int32_t), allocatable :: atype(:), kind(:)
     integer(c_int32_t), allocatable :: cend(:)   ! last atom of every chain (molecule copy), md_run's ixcend
     integer(c_int32_t), allocatable :: excl(:,:), excess(:,:)
     integer(c_int32_t), allocatable :: snat(:), satm(:,:)
     real(8), allocatable :: sdist(:,:)
     integer(c_int32_t), allocatable :: batm(:,:), aatm(:,:), tatm(:,:), tdiv(:), patm(:,:)
     real(8), allocatable :: bfc(:), br0(:), afc(:), ath(:), tfc(:), trot(:), tph(:), p14v(:), p14e(:)
  end type top_t

contains

  ! One logical record: continuation lines joined, comment stripped, split
  ! into tokens.  Returns .false. at end of file.
  logical function next_record(u, tok, ntok, sect)
    integer, intent(in) :: u
    character(len=toklen), intent(out) :: tok(maxtok)
    integer, intent(out) :: ntok
    character(len=*), intent(out) :: sect
    character(len=4096) :: buf
    character(len=256) :: line
    integer :: ios, p, q, l
    next_record = .false.
    ntok = 0
    sect = ' '
    buf = ' '
    do
       read(u, '(a)', iostat=ios) line
       if (ios /= 0) return
       p = index(line, ';')
       if (p > 0) line(p:) = ' '
       if (index(line, 'TPL>') > 0 .or. index(line, 'SHAKE>') > 0) then
          sect = adjustl(line(index(line, '>') + 1:))
          next_record = .true.
          return
       end if
       l = len_trim(line)
       if (l >= 2) then
          if (line(l - 1:l) == '->') then
             buf = trim(buf) // ' ' // line(1:l - 2)
             cycle
          end if
       end if
       buf = trim(buf) // ' ' // line
       if (len_trim(buf) == 0) cycle
       exit
    end do
    ! split
    p = 1
    l = len_trim(buf)
    do while (p <= l)
       do while (p <= l)
          if (buf(p:p) /= ' ') exit
          p = p + 1
       end do
       if (p > l) exit
       q = p
       do while (q <= l)
          if (buf(q:q) == ' ') exit
          q = q + 1
       end do
       ntok = ntok + 1
       tok(ntok) = buf(p:q - 1)
       p = q
    end do
    next_record = .true.
  end function next_record

  integer function ival(s)
    character(len=*), intent(in) :: s
    read(s, *) ival
  end function ival

  real(8) function rval(s)
    character(len=*), intent(in) :: s
    read(s, *) rval
  end function rval

  integer function find_mol(mols, nm, name)
    type(mol_t), intent(in) :: mols(:)
    integer, intent(in) :: nm
    character(len=*), intent(in) :: name
    integer :: k
    find_mol = 0
    do k = 1, nm
       if (trim(mols(k)%name) == trim(name)) then
          find_mol = k
          return
       end if
    end do
  end function find_mol

  ! TPL + SHK -> expanded tables.  ier: 0, or 1 (file), 2 (format).
  subroutine read_topology(tplpath, shkpath, diel, top, ier)
    character(len=*), intent(in) :: tplpath, shkpath
    real(8), intent(in) :: diel
    type(top_t), intent(out) :: top
    integer, intent(out) :: ier
    type(mol_t), allocatable :: mols(:)
    character(len=toklen) :: tok(maxtok)
    character(len=64) :: sect, cur
    integer :: u, ntok, nm, k, im, ios
    real(8), allocatable :: tr2(:), te(:), tv(:), ts(:)
    ier = 0
    open(newunit=u, file=tplpath, status='old', action='read', iostat=ios)
    if (ios /= 0) then
       ier = 1
       return
    end if
    allocate(mols(64))
    nm = 0
    cur = ' '
    im = 0
    allocate(tr2(256), te(256), tv(256), ts(256))
    top%ntyp = 0
    do while (next_record(u, tok, ntok, sect))
       if (len_trim(sect) > 0) then
          cur = sect
          im = -1            ! next record names the molecule (per-molecule sections)
          cycle
       end if
       select case (trim(cur))
       case ('MOLECULES')
          nm = nm + 1
          mols(nm)%name = tok(1)
          mols(nm)%ncopy = ival(tok(2))
       case ('ATOMS', 'BONDS', 'ANGLES', 'TORSIONS')
          if (im == -1) then
             im = find_mol(mols, nm, tok(1))
             if (im == 0) then
                ier = 2
                return
             end if
             cycle
          end if
          call add_term(mols(im), trim(cur), tok, ntok, ier)
          if (ier /= 0) return
       case ('NONBONDS')
          k = ival(tok(1))
          if (k > size(tr2)) then
             ier = 2
             return
          end if
          top%ntyp = max(top%ntyp, k)
          tr2(k) = rval(tok(4)); te(k) = rval(tok(5))
          ts(k) = rval(tok(6)); tv(k) = rval(tok(7))   ! 1-4 elec scale, 1-4 vdW scale
       end select
    end do
    close(u)
    if (len_trim(shkpath) > 0) then
       open(newunit=u, file=shkpath, status='old', action='read', iostat=ios)
       if (ios /= 0) then
          ier = 1
          return
       end if
       im = 0
       do while (next_record(u, tok, ntok, sect))
          if (len_trim(sect) > 0) then
             im = -1
             cycle
          end if
          if (im == -1) then
             im = find_mol(mols, nm, tok(1))
             if (im == 0) then
                ier = 2
                return
             end if
             cycle
          end if
          call add_term(mols(im), 'SHAKE', tok, ntok, ier)
          if (ier /= 0) return
       end do
       close(u)
    end if
    call expand(mols(1:nm), top, diel, tr2, te, ts, tv, ier)
  end subroutine read_topology

  subroutine grow_i(a, n2)
    integer, allocatable, intent(inout) :: a(:,:)
    integer, intent(in) :: n2
    integer, allocatable :: b(:,:)
    if (.not. allocated(a)) then
       allocate(a(n2, 16))
       return
    end if
    allocate(b(size(a, 1), 2 * size(a, 2)))
    b(:, 1:size(a, 2)) = a
    call move_alloc(b, a)
  end subroutine grow_i

  subroutine grow_r(a, n2)
    real(8), allocatable, intent(inout) :: a(:,:)
    integer, intent(in) :: n2
    real(8), allocatable :: b(:,:)
    if (.not. allocated(a)) then
       allocate(a(n2, 16))
       return
    end if
    allocate(b(size(a, 1), 2 * size(a, 2)))
    b(:, 1:size(a, 2)) = a
    call move_alloc(b, a)
  end subroutine grow_r

  subroutine push_r(a, n, v)
    real(8), allocatable, intent(inout) :: a(:)
    integer, intent(in) :: n
    real(8), intent(in) :: v
    real(8), allocatable :: b(:)
    if (.not. allocated(a)) allocate(a(16))
    if (n > size(a)) then
       allocate(b(2 * size(a)))
       b(1:size(a)) = a
       call move_alloc(b, a)
    end if
    a(n) = v
  end subroutine push_r

  subroutine push_i(a, n, v)
    integer, allocatable, intent(inout) :: a(:)
    integer, intent(in) :: n, v
    integer, allocatable :: b(:)
    if (.not. allocated(a)) allocate(a(16))
    if (n > size(a)) then
       allocate(b(2 * size(a)))
       b(1:size(a)) = a
       call move_alloc(b, a)
    end if
    a(n) = v
  end subroutine push_i

  ! one record of a per-molecule section
  subroutine add_term(m, what, tok, ntok, ier)
    type(mol_t), intent(inout) :: m
    character(len=*), intent(in) :: what
    character(len=toklen), intent(in) :: tok(:)
    integer, intent(in) :: ntok
    integer, intent(inout) :: ier
    integer :: k, np, n
    select case (what)
    case ('ATOMS')
       ! name, atom type name, type, residue, residue no, mass, Rmin/2, charge,
       ! forward 1-2/1-3/1-4 partner counts, the partner offsets, 7 more
       m%na = m%na + 1
       n = m%na
       call push_r(m%mass, n, rval(tok(6)))
       call push_r(m%chg, n, rval(tok(8)))
       call push_i(m%typ, n, ival(tok(3)))
       np = ival(tok(9)) + ival(tok(10)) + ival(tok(11))
       call push_i(m%npart, n, np)
       if (.not. allocated(m%part)) allocate(m%part(32, 16))
       if (n > size(m%part, 2)) call grow_i(m%part, 32)
       if (np > 32 .or. ntok < 11 + np) then
          ier = 2
          return
       end if
       do k = 1, np
          m%part(k, n) = ival(tok(11 + k))
       end do
    case ('BONDS')
       m%nb = m%nb + 1
       if (.not. allocated(m%b)) allocate(m%b(2, 16))
       if (m%nb > size(m%b, 2)) call grow_i(m%b, 2)
       m%b(1, m%nb) = ival(tok(1)); m%b(2, m%nb) = ival(tok(2))
       call push_r(m%bfc, m%nb, rval(tok(3)))
       call push_r(m%br0, m%nb, rval(tok(4)))
    case ('ANGLES')
       m%nang = m%nang + 1
       if (.not. allocated(m%a)) allocate(m%a(3, 16))
       if (m%nang > size(m%a, 2)) call grow_i(m%a, 3)
       m%a(1:3, m%nang) = [ival(tok(1)), ival(tok(2)), ival(tok(3))]
       call push_r(m%afc, m%nang, rval(tok(4)))
       call push_r(m%ath, m%nang, rval(tok(5)))
    case ('TORSIONS')
       ! i j k l, fc, division, periodicity, phase (deg), 1-4 flag
       m%nt = m%nt + 1
       if (.not. allocated(m%t)) allocate(m%t(4, 16))
       if (m%nt > size(m%t, 2)) call grow_i(m%t, 4)
       m%t(1:4, m%nt) = [ival(tok(1)), ival(tok(2)), ival(tok(3)), ival(tok(4))]
       call push_r(m%tfc, m%nt, rval(tok(5)))
       call push_i(m%tdiv, m%nt, ival(tok(6)))
       call push_r(m%trot, m%nt, rval(tok(7)))
       call push_r(m%tph, m%nt, rval(tok(8)))
       call push_i(m%tnbf, m%nt, ival(tok(9)))
    case ('SHAKE')
       ! number of atoms, the atoms, the distances of pairs (1,2) (2,3) (3,1) ...
       m%ns = m%ns + 1
       n = ival(tok(1))
       if (n < 2 .or. n > 4 .or. ntok < 1 + n + n * (n - 1) / 2) then
          ier = 2
          return
       end if
       call push_i(m%sn, m%ns, n)
       if (.not. allocated(m%sa)) allocate(m%sa(4, 16), m%sd(6, 16))
       if (m%ns > size(m%sa, 2)) then
          call grow_i(m%sa, 4)
          call grow_r(m%sd, 6)
       end if
       m%sa(:, m%ns) = 0
       m%sd(:, m%ns) = 0.0d0
       do k = 1, n
          m%sa(k, m%ns) = ival(tok(1 + k))
       end do
       do k = 1, n * (n - 1) / 2
          m%sd(k, m%ns) = rval(tok(1 + n + k))
       end do
    end select
  end subroutine add_term

  ! molecules x copies -> atom tables in jp_sys_desc's form
  subroutine expand(mols, top, diel, tr2, te, ts, tv, ier)
    type(mol_t), intent(in) :: mols(:)
    type(top_t), intent(inout) :: top
    real(8), intent(in) :: diel, tr2(:), te(:), ts(:), tv(:)
    integer, intent(inout) :: ier
    integer :: im, c, off, a, k, n, nb, na, nt, ns, nx, ne, g, sz, q, i1, i4, nch
    integer, allocatable :: xp(:,:), gorder(:), nbf(:)
    logical, allocatable :: offb(:), offa(:)
    integer :: np14
    real(8) :: ccoef
    integer(c_int32_t), allocatable :: bt(:,:), at(:,:)
    real(8), allocatable :: bfc(:), br0(:), afc(:), ath(:)
    n = 0; nb = 0; na = 0; nt = 0; ns = 0; nx = 0
    do im = 1, size(mols)
       n = n + mols(im)%na * mols(im)%ncopy
       nb = nb + mols(im)%nb * mols(im)%ncopy
       na = na + mols(im)%nang * mols(im)%ncopy
       nt = nt + mols(im)%nt * mols(im)%ncopy
       ns = ns + mols(im)%ns * mols(im)%ncopy
       if (mols(im)%na > 0) nx = nx + sum(mols(im)%npart(1:mols(im)%na)) * mols(im)%ncopy
    end do
    top%n = n
    allocate(top%mass(n), top%chg(n), top%atype(n), top%kind(n))
    allocate(top%cend(max(sum(mols(:)%ncopy), 1)))
    top%cend = 0
    nch = 0
    allocate(xp(2, max(nx, 1)), bt(2, max(nb, 1)), bfc(max(nb, 1)), br0(max(nb, 1)))
    allocate(at(3, max(na, 1)), afc(max(na, 1)), ath(max(na, 1)))
    allocate(top%tatm(4, max(nt, 1)), top%tfc(max(nt, 1)), top%trot(max(nt, 1)), top%tph(max(nt, 1)), &
             top%tdiv(max(nt, 1)))
    allocate(top%snat(max(ns, 1)), top%satm(4, max(ns, 1)), top%sdist(6, max(ns, 1)), gorder(max(ns, 1)))
    off = 0; nx = 0; nb = 0; na = 0; nt = 0; ns = 0
    do im = 1, size(mols)
       associate (m => mols(im))
         do c = 1, m%ncopy
            do a = 1, m%na
               top%mass(off + a) = m%mass(a)
               top%chg(off + a) = m%chg(a)
               top%atype(off + a) = m%typ(a)
               top%kind(off + a) = im
               do k = 1, m%npart(a)
                  nx = nx + 1
                  xp(1, nx) = min(off + a, off + a + m%part(k, a))
                  xp(2, nx) = max(off + a, off + a + m%part(k, a))
               end do
            end do
            do k = 1, m%nb
               nb = nb + 1
               bt(:, nb) = off + m%b(:, k); bfc(nb) = m%bfc(k); br0(nb) = m%br0(k)
            end do
            do k = 1, m%nang
               na = na + 1
               at(:, na) = off + m%a(:, k); afc(na) = m%afc(k); ath(na) = m%ath(k)
            end do
            do k = 1, m%nt
               nt = nt + 1
               top%tatm(:, nt) = off + m%t(:, k); top%tfc(nt) = m%tfc(k); top%tdiv(nt) = m%tdiv(k)
               top%trot(nt) = m%trot(k); top%tph(nt) = m%tph(k)
            end do
            do k = 1, m%ns
               ns = ns + 1
               top%snat(ns) = m%sn(k)
               top%satm(:, ns) = 0
               top%satm(1:m%sn(k), ns) = off + m%sa(1:m%sn(k), k)
               top%sdist(:, ns) = m%sd(:, k) * m%sd(:, k)
            end do
            off = off + m%na
            nch = nch + 1                 ! inptpl.f90:279, 680: ixcend(chain) = atoms so far
            top%cend(nch) = off
         end do
       end associate
    end do
    ! SHAKE groups by size, input order within a size (setshk.f90 modshk)
    q = 0
    do sz = 2, 4
       do g = 1, ns
          if (top%snat(g) == sz) then
             q = q + 1
             gorder(q) = g
          end if
       end do
    end do
    top%snat(1:ns) = top%snat(gorder(1:ns))
    top%satm(:, 1:ns) = top%satm(:, gorder(1:ns))
    top%sdist(:, 1:ns) = top%sdist(:, gorder(1:ns))
    ! unique excluded pairs in ascending order
    call sort_pairs(xp, nx, top%excl, top%nexcl)
    top%nshake = ns
    ! excess pairs: bonds, angles' 1-3, torsions' 1-4 with the flag
    ne = nb + na
    do im = 1, size(mols)
       if (mols(im)%nt > 0) ne = ne + count(mols(im)%tnbf(1:mols(im)%nt) == 1) * mols(im)%ncopy
    end do
    allocate(top%excess(2, max(ne, 1)))
    top%excess(:, 1:nb) = bt(:, 1:nb)
    do k = 1, na
       top%excess(1, nb + k) = at(1, k); top%excess(2, nb + k) = at(3, k)
    end do
    ne = nb + na
    np14 = 0
    call flags14(mols, nt)
    do k = 1, nt
       if (nbf(k) == 1) then
          ne = ne + 1
          top%excess(1, ne) = top%tatm(1, k); top%excess(2, ne) = top%tatm(4, k)
          np14 = np14 + 1
       end if
    end do
    top%nexcess = ne
    ! bonds and angles setshk keeps
    call shake_removed(bt, nb, at, na, top%snat, top%satm, ns, offb, offa)
    ! kept terms; a set whose force constants are all zero contributes nothing
    allocate(top%batm(2, max(nb, 1)), top%bfc(max(nb, 1)), top%br0(max(nb, 1)))
    k = 0
    do q = 1, nb
       if (offb(q)) cycle
       k = k + 1
       top%batm(:, k) = bt(:, q); top%bfc(k) = bfc(q); top%br0(k) = br0(q)
    end do
    if (all(top%bfc(1:k) == 0.0d0)) k = 0
    top%nbond = k
    allocate(top%aatm(3, max(na, 1)), top%afc(max(na, 1)), top%ath(max(na, 1)))
    k = 0
    do q = 1, na
       if (offa(q)) cycle
       k = k + 1
       top%aatm(:, k) = at(:, q); top%afc(k) = afc(q)
       top%ath(k) = ath(q) * (3.141592653589793d0 / 180.0d0)   ! inptpl.f90 rdangl
    end do
    if (all(top%afc(1:k) == 0.0d0)) k = 0
    top%nangle = k
    ! types and the 1-4 pairs (AMBER scales: the smaller of the two types')
    allocate(top%rmin2(top%ntyp), top%eps(top%ntyp), top%s14e(top%ntyp), top%s14v(top%ntyp))
    top%rmin2 = tr2(1:top%ntyp); top%eps = te(1:top%ntyp)
    top%s14e = ts(1:top%ntyp); top%s14v = tv(1:top%ntyp)
    ccoef = 332.06378d0 / diel
    allocate(top%patm(2, max(np14, 1)), top%p14v(max(np14, 1)), top%p14e(max(np14, 1)))
    np14 = 0
    do k = 1, nt
       if (nbf(k) /= 1) cycle
       np14 = np14 + 1
       i1 = top%tatm(1, k); i4 = top%tatm(4, k)
       top%patm(:, np14) = [i1, i4]
       top%p14v(np14) = min(top%s14v(top%atype(i1)), top%s14v(top%atype(i4))) * nbf(k)
       top%p14e(np14) = min(top%s14e(top%atype(i1)), top%s14e(top%atype(i4))) * nbf(k) * ccoef &
                        * top%chg(i1) * top%chg(i4)
    end do
    top%n14 = np14
    call keep_torsions(top, nt)
  contains
    subroutine flags14(ms, ntt)
      type(mol_t), intent(in) :: ms(:)
      integer, intent(in) :: ntt
      integer :: i, cc, kk, q2
      allocate(nbf(max(ntt, 1)))
      q2 = 0
      do i = 1, size(ms)
         do cc = 1, ms(i)%ncopy
            do kk = 1, ms(i)%nt
               q2 = q2 + 1
               nbf(q2) = ms(i)%tnbf(kk)
            end do
         end do
      end do
    end subroutine flags14
  end subroutine expand

  ! torsions with |fc| >= epsilon (incel.f90:993-1009), phase in radians
  subroutine keep_torsions(top, nt)
    type(top_t), intent(inout) :: top
    integer, intent(in) :: nt
    integer :: k, q
    q = 0
    do k = 1, nt
       if (abs(top%tfc(k)) < epsilon(1.0d0)) cycle
       q = q + 1
       top%tatm(:, q) = top%tatm(:, k); top%tfc(q) = top%tfc(k); top%tdiv(q) = top%tdiv(k)
       top%trot(q) = top%trot(k); top%tph(q) = top%tph(k) * (3.141592653589793d0 / 180.0d0)
    end do
    top%ntors = q
  end subroutine keep_torsions

  ! unique (i < j) pairs in ascending (i, j) order
  subroutine sort_pairs(xp, nx, out, q)
    integer, intent(in) :: xp(:,:)
    integer, intent(in) :: nx
    integer(c_int32_t), allocatable, intent(out) :: out(:,:)
    integer, intent(out) :: q
    integer(8), allocatable :: key(:)
    integer(8), parameter :: big = 2147483647_8
    integer :: k
    allocate(key(max(nx, 1)))
    do k = 1, nx
       key(k) = int(xp(1, k), 8) * big + int(xp(2, k), 8)
    end do
    call isort(key, nx)
    q = 0
    do k = 1, nx
       if (k > 1) then
          if (key(k) == key(k - 1)) cycle
       end if
       q = q + 1
       key(q) = key(k)
    end do
    allocate(out(2, max(q, 1)))
    out = 0
    do k = 1, q
       out(1, k) = int(key(k) / big, c_int32_t)
       out(2, k) = int(mod(key(k), big), c_int32_t)
    end do
  end subroutine sort_pairs

  ! heap sort of n 64-bit keys
  subroutine isort(a, n)
    integer(8), intent(inout) :: a(:)
    integer, intent(in) :: n
    integer :: i, j, k, e
    integer(8) :: t
    do i = n / 2, 1, -1
       call sift(i, n)
    end do
    do e = n, 2, -1
       t = a(1); a(1) = a(e); a(e) = t
       call sift(1, e - 1)
    end do
  contains
    subroutine sift(s, m)
      integer, intent(in) :: s, m
      integer :: r, c
      integer(8) :: v
      r = s
      v = a(r)
      do
         c = 2 * r
         if (c > m) exit
         if (c < m) then
            if (a(c + 1) > a(c)) c = c + 1
         end if
         if (a(c) <= v) exit
         a(r) = a(c)
         r = c
      end do
      a(r) = v
    end subroutine sift
  end subroutine isort

  ! setshk.f90 reabnd / reaang: the bonds (1-2) and angles (1-3) a SHAKE
  ! constraint holds; per group size a forward scan from after the last match,
  ! abandoned once the term's atoms pass the constraint's
  subroutine shake_removed(bt, nb, at, na, snat, satm, ns, offb, offa)
    integer(c_int32_t), intent(in) :: bt(:,:), at(:,:), snat(:), satm(:,:)
    integer, intent(in) :: nb, na, ns
    logical, allocatable, intent(out) :: offb(:), offa(:)
    integer, parameter :: pl(2, 6) = reshape([1, 2, 2, 3, 3, 1, 4, 1, 3, 4, 2, 4], [2, 6])
    integer :: sz, g, p, nst, ib, i1, i2, j1, j2
    allocate(offb(max(nb, 1)), offa(max(na, 1)))
    offb = .false.; offa = .false.
    do sz = 2, 4
       nst = 1
       do g = 1, ns
          if (snat(g) /= sz) cycle
          do p = 1, sz * (sz - 1) / 2
             i1 = satm(pl(1, p), g); i2 = satm(pl(2, p), g)
             do ib = nst, nb
                j1 = bt(1, ib); j2 = bt(2, ib)
                if ((i1 == j1 .and. i2 == j2) .or. (i1 == j2 .and. i2 == j1)) then
                   offb(ib) = .true.
                   if (sz /= 3) then
                      nst = ib + 1
                   else
                      nst = max(ib - 1, 1)
                   end if
                   exit
                end if
                if (passed(i1, i2, j1, j2)) exit
             end do
          end do
       end do
    end do
    do sz = 3, 4
       nst = 1
       do g = 1, ns
          if (snat(g) /= sz) cycle
          do p = 1, sz * (sz - 1) / 2
             i1 = satm(pl(1, p), g); i2 = satm(pl(2, p), g)
             do ib = nst, na
                j1 = at(1, ib); j2 = at(3, ib)
                if ((i1 == j1 .and. i2 == j2) .or. (i1 == j2 .and. i2 == j1)) then
                   offa(ib) = .true.
                   if (sz == 3) then
                      nst = max(ib - 1, 1)
                   else
                      nst = max(ib - 2, 1)
                   end if
                   exit
                end if
                if (passed(i1, i2, j1, j2)) exit
             end do
          end do
       end do
    end do
  contains
    logical function passed(a1, a2, b1, b2)
      integer, intent(in) :: a1, a2, b1, b2
      passed = (a1 < b1 .and. a2 < b2) .or. (a1 < b2 .and. a2 < b1)
    end function passed
  end subroutine shake_removed

  ! BINA coordinates (inpcrb.f90): a title record, the atom count, x, y, z
  subroutine read_crd_bina(path, n, x, ier)
    character(len=*), intent(in) :: path
    integer, intent(in) :: n
    real(8), intent(out) :: x(3, n)
    integer, intent(out) :: ier
    character(len=1) :: title
    integer(c_int32_t) :: m
    real(8), allocatable :: c(:)
    integer :: u, ios, k
    ier = 0
    open(newunit=u, file=path, form='unformatted', access='sequential', status='old', action='read', &
         iostat=ios)
    if (ios /= 0) then
       ier = 1
       return
    end if
    read(u) title
    read(u) m
    if (m /= n) then
       ier = 2
       close(u)
       return
    end if
    allocate(c(n))
    do k = 1, 3
       read(u) c
       x(k, :) = c
    end do
    close(u)
  end subroutine read_crd_bina

  ! PDB coordinates (inppdb.f90): every ATOM / HETATM record, x y z as
  ! 3f8.3 after column 30.  ier: 0, 1 (file), 2 (atom count), 3 (format)
  subroutine read_crd_pdb(path, n, x, ier)
    character(len=*), intent(in) :: path
    integer, intent(in) :: n
    real(8), intent(out) :: x(3, n)
    integer, intent(out) :: ier
    character(len=80) :: line
    integer :: u, ios, m
    ier = 0
    open(newunit=u, file=path, status='old', action='read', iostat=ios)
    if (ios /= 0) then
       ier = 1
       return
    end if
    m = 0
    do
       read(u, '(a80)', iostat=ios) line
       if (ios /= 0) exit
       if (line(1:4) /= 'ATOM' .and. line(1:4) /= 'HETA') cycle
       m = m + 1
       if (m > n) exit
       read(line, '(30x,3f8.3)', iostat=ios) x(1:3, m)
       if (ios /= 0) then
          ier = 3
          exit
       end if
    end do
    close(u)
    if (ier == 0 .and. m /= n) ier = 2
  end subroutine read_crd_pdb

end module jp_input
