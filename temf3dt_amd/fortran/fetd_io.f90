! fetd_io.f90 -- host-side file formats of the FETD driver: MEDIT .mesh and
! .cntl readers, .field/.field2 writers.
!
! Mirrors the reference's input/output code (yxddyxzh2016/TEMF3DT):
!   read_medit  -- model_mesh reader, src/FETD_FWD.f90:126-167 (sections keyed
!                  by the first five characters: Verti/Edges/Trian/Tetra; Edges
!                  with a non-zero physical tag are the source wire)
!   read_cntl   -- model_ctl keys, src/FETD_FWD.f90:284-497, with the line
!                  grammar of parseCode, src/FETD_UTILITY.f90:678-731
!   field rows  -- solution_output, src/FETD_FWD.f90:1602-1641: format
!                  (i6,e15.5,2x,R(m e15.5,2x)); a gate is filled from the first
!                  row with t >= gate by linear interpolation with the row before
!   msh slices  -- msh_output, src/FETD_FWD.f90:1713-1864: a copy of ROOT.msh
!                  with one $ElementData section per selected component
!   rescaling   -- the small-loop variant, src_small/FETD_FWD.f90:314-328,
!                  1526,1540
module fetd_io
  use iso_c_binding, only: c_int, c_double
  implicit none
  private
  public :: cntl_t, read_medit, read_cntl, tet_materials, tet_tensors, aniso_tensor, field_writer_t, fw_open, &
            fw_row, fw_close, rescale_domains, msh_tet_count, write_msh_slice

  real(8), parameter :: mu0 = 3.14159265358979323846264338327950288419716939937510d0 * 4.0d-7

  type cntl_t
    integer :: rec_num = 0, dom_num = 0, src_num = 0, bound_cond = 0
    real(8), allocatable :: rec(:, :)        ! (3, rec_num)
    real(8), allocatable :: dom(:, :)        ! (4, dom_num): id, sigma, permittivity, permeability
    logical :: aniso = .false.
    real(8), allocatable :: adom(:, :)       ! (19, dom_num): id, then (3 values, 3 angles) for sigma, eps, mu
    integer :: source_type = 1, pulse_type = 1
    real(8) :: pulse_width = 0.0d0, current = 1.0d0
    integer :: sp_mthd = 0, sp_division = 0, sp_double = 0, sp_dblesize = 0
    real(8) :: time_maximum = 0.0d0, rec_charac = 10.0d0
    integer :: o_f(9) = 0
    real(8), allocatable :: time_points(:)
    integer :: vtk_num = 0
    real(8), allocatable :: vtk_times(:)
    real(8) :: sig_parameter = 1.0d0     ! small-loop variant's scale (1 = off)
  end type cntl_t

  type field_writer_t
    integer :: u1 = -1, u2 = -1, nrec = 0, nsel = 0, nrow = 0
    integer :: o_v(9) = 0
    character(64) :: fmt
    real(8), allocatable :: gates(:)
    logical, allocatable :: gate_ok(:)
    real(8), allocatable :: prev(:, :)       ! (9, nrec) previous row
    real(8) :: tprev = 0.0d0
    logical :: rescale = .false.             ! small-loop variant: written v / sp**2 * pi**log10(sp)
    real(8) :: sp2 = 1.0d0, spk = 1.0d0
  end type field_writer_t

  real(8), parameter :: pi = 3.14159265358979323846264338327950288419716939937510d0

contains

  ! ---- MEDIT --------------------------------------------------------------
  subroutine read_medit(path, nN, xyz, nT, tet, tet_phy, nseg, seg, ierr, msg)
    character(*), intent(in) :: path
    integer, intent(out) :: nN, nT, nseg, ierr
    real(c_double), allocatable, intent(out) :: xyz(:, :)   ! (3, nN)
    integer(c_int), allocatable, intent(out) :: tet(:, :)    ! (4, nT)
    integer(c_int), allocatable, intent(out) :: tet_phy(:)
    integer(c_int), allocatable, intent(out) :: seg(:, :)    ! (2, nseg)
    character(*), intent(out) :: msg
    integer :: u, ios, i, n, e(3), phys
    integer, allocatable :: etmp(:, :)
    character(256) :: line
    character(5) :: key
    real(8) :: ref
    integer :: t5(5), tri(4)

    nN = 0; nT = 0; nseg = 0; ierr = 0; msg = ' '
    open(newunit=u, file=path, status='old', action='read', iostat=ios)
    if (ios /= 0) then
      ierr = 1; msg = 'cannot open '//trim(path); return
    end if
    do
      read(u, '(a)', iostat=ios) line
      if (ios /= 0) exit
      key = adjustl(line)
      select case (key)
      case ('Verti')
        read(u, *) n
        nN = n
        allocate(xyz(3, n))
        do i = 1, n
          read(u, *) xyz(1, i), xyz(2, i), xyz(3, i), ref
        end do
      case ('Edges')
        read(u, *) n
        allocate(etmp(2, n))
        nseg = 0
        do i = 1, n
          read(u, *) e(1), e(2), phys
          if (phys /= 0) then
            nseg = nseg + 1
            etmp(:, nseg) = e(1:2)
          end if
        end do
        allocate(seg(2, nseg))
        seg(:, 1:nseg) = etmp(:, 1:nseg)
        deallocate(etmp)
      case ('Trian')
        read(u, *) n
        do i = 1, n
          read(u, *) tri
        end do
      case ('Tetra')
        read(u, *) n
        nT = n
        allocate(tet(4, n), tet_phy(n))
        do i = 1, n
          read(u, *) t5
          tet(:, i) = t5(1:4)
          tet_phy(i) = t5(5)
        end do
      end select
    end do
    close(u)
    if (nN == 0 .or. nT == 0) then
      ierr = 2; msg = trim(path)//': mesh has no Vertices or Tetrahedra'
    end if
    if (.not. allocated(seg)) allocate(seg(2, 0))
  end subroutine read_medit

  ! ---- .cntl --------------------------------------------------------------
  ! One line -> lower-case key and value (comment lines, blank lines and lines
  ! without a colon give an empty key).
  subroutine split_code(line, key, val)
    character(*), intent(inout) :: line
    character(*), intent(out) :: key, val
    integer :: i, i0, ic, ip
    key = ' '; val = ' '
    do i = 1, len(line)
      if (ichar(line(i:i)) == 9) line(i:i) = ' '
    end do
    i0 = verify(line, ' ')
    if (i0 == 0 .or. i0 >= len(line)) return
    if (line(i0:i0) == '%' .or. line(i0:i0) == '!') return
    ic = index(line, ':')
    if (ic - 1 < i0) return
    key = line(i0:ic - 1)
    do i = 1, len_trim(key)
      if (key(i:i) >= 'A' .and. key(i:i) <= 'Z') key(i:i) = achar(iachar(key(i:i)) + 32)
    end do
    val = adjustl(line(ic + 1:))
    ip = index(val, '%')
    if (ip > 0 .and. ip < len_trim(val)) val(ip:) = ' '
    ip = index(val, '!')
    if (ip > 0 .and. ip < len_trim(val)) val(ip:) = ' '
  end subroutine split_code

  subroutine read_cntl(path, c, ierr, msg)
    character(*), intent(in) :: path
    type(cntl_t), intent(out) :: c
    integer, intent(out) :: ierr
    character(*), intent(out) :: msg
    integer :: u, ios, i, n
    character(512) :: line
    character(128) :: key
    character(512) :: val

    ierr = 0; msg = ' '
    open(newunit=u, file=path, status='old', action='read', iostat=ios)
    if (ios /= 0) then
      ierr = 1; msg = 'cannot open '//trim(path); return
    end if
    do
      read(u, '(a)', iostat=ios) line
      if (ios /= 0) exit
      call split_code(line, key, val)
      if (len_trim(key) == 0) cycle
      select case (trim(key))
      case ('receivers')
        read(val, *) n
        c%rec_num = n
        allocate(c%rec(3, n))
        do i = 1, n
          read(u, *) c%rec(:, i)
        end do
      case ('isotropic')
        read(val, *) n
        c%dom_num = n
        allocate(c%dom(4, n))
        do i = 1, n
          read(u, *) c%dom(:, i)
        end do
      case ('anisotropic')
        ! three records per domain: "domain v1 v2 v3 angle1 angle2 angle3" for
        ! conductivity, permittivity, permeability (src/FETD_FWD.f90:368-377)
        read(val, *) n
        c%dom_num = n
        c%aniso = .true.
        allocate(c%adom(19, n), c%dom(4, n))
        do i = 1, n
          read(u, *) c%adom(1, i), c%adom(2:7, i)
          read(u, *) c%adom(1, i), c%adom(8:13, i)
          read(u, *) c%adom(1, i), c%adom(14:19, i)
          c%dom(:, i) = [c%adom(1, i), c%adom(2, i), c%adom(8, i), c%adom(14, i)]
        end do
      case ('boundary condition')
        read(val, *) c%bound_cond
      case ('source')
        read(val, *) c%src_num
        read(u, *) c%source_type, c%pulse_type, c%pulse_width, c%current
      case ('sp_mthd')
        read(val, *) c%sp_mthd
      case ('sp_division')
        read(val, *) c%sp_division
      case ('time_maximum')
        read(val, *) c%time_maximum
      case ('sp_double')
        read(val, *) c%sp_double
      case ('sp_dblesize')
        read(val, *) c%sp_dblesize
      case ('output')
        read(u, *) c%o_f
      case ('vtkout')
        read(val, *) c%vtk_num
        allocate(c%vtk_times(max(c%vtk_num, 0)))
        do i = 1, c%vtk_num
          read(u, *) c%vtk_times(i)
        end do
      case ('time_points')
        read(val, *) n
        allocate(c%time_points(n))
        do i = 1, n
          read(u, *) c%time_points(i)
        end do
      case ('rec_charac')
        read(val, *) c%rec_charac
      end select
    end do
    close(u)
    if (ierr /= 0) return
    if (.not. allocated(c%time_points)) allocate(c%time_points(0))
    if (.not. allocated(c%vtk_times)) allocate(c%vtk_times(0))
    ! required keys (src/FETD_FWD.f90:490-497)
    if (c%rec_num == 0) then
      ierr = 4; msg = 'cmod%rec_num error'
    else if (c%dom_num == 0) then
      ierr = 4; msg = 'cmod%dom_num error'
    else if (c%src_num == 0) then
      ierr = 4; msg = 'cmod%src_num error'
    else if (c%sp_mthd == 0) then
      ierr = 4; msg = 'cmod%step_mthd error'
    else if (c%sp_division == 0) then
      ierr = 4; msg = 'cmod%step_dvsi error'
    else if (abs(c%time_maximum) < 1.0d-20) then
      ierr = 4; msg = 'cmod%time_maxm error'
    else if (c%sp_double == 0) then
      ierr = 4; msg = 'cmod%step_dble error'
    else if (count(c%o_f == 1) == 0) then
      ierr = 4; msg = 'cmod%o_f error'
    end if
  end subroutine read_cntl

  ! The small-loop variant (src_small/FETD_FWD.f90:314-328): sp = 1 / sigma
  ! of the last domain when there are several (the branch written to take
  ! the smallest is unreachable); sigma of domains 2..n times sp.
  subroutine rescale_domains(c)
    type(cntl_t), intent(inout) :: c
    integer :: i
    real(8) :: sp
    sp = 1.0d0
    do i = 2, c%dom_num
      sp = c%dom(2, i)
    end do
    sp = 1.0d0 / sp
    do i = 2, c%dom_num
      c%dom(2, i) = c%dom(2, i) * sp
    end do
    c%sig_parameter = sp
  end subroutine rescale_domains

  ! per-tet sigma and 1/mu from the isotropic domain table (src/FETD_FWD.f90:347-364)
  subroutine tet_materials(c, nT, tet_phy, sigma, inv_mu, ierr)
    type(cntl_t), intent(in) :: c
    integer, intent(in) :: nT
    integer(c_int), intent(in) :: tet_phy(nT)
    real(c_double), intent(out) :: sigma(nT), inv_mu(nT)
    integer, intent(out) :: ierr
    integer :: i, d
    logical :: found
    ierr = 0
    do i = 1, nT
      found = .false.
      do d = 1, c%dom_num
        if (tet_phy(i) == nint(c%dom(1, d))) then
          sigma(i) = c%dom(2, d)
          inv_mu(i) = 1.0d0 / (c%dom(4, d) * mu0)
          found = .true.
          exit
        end if
      end do
      if (.not. found) then
        ierr = 1; return
      end if
    end do
  end subroutine tet_materials

  ! R diag(v) R^T with R = Rx Ry Rz from three angles in degrees, times
  ! scale; as in the reference the (2,2) entry of Rz takes the cosine of the
  ! second angle (src/FETD_FWD.f90:385-392).  A is row-major in a(1:9).
  subroutine aniso_tensor(v, ang, scale, a)
    real(8), intent(in) :: v(3), ang(3), scale
    real(8), intent(out) :: a(9)
    real(8), parameter :: d2r = 3.141592653589793238462643383279502884197d0 / 180.0d0
    real(8) :: dx(3, 3), dy(3, 3), dz(3, 3), t(3, 3), dm(3, 3), w(3, 3), a1, a2, a3, acc
    integer :: i, j, k
    a1 = ang(1) * d2r; a2 = ang(2) * d2r; a3 = ang(3) * d2r
    dx = reshape([1.0d0, 0.0d0, 0.0d0, 0.0d0, cos(a1), sin(a1), 0.0d0, -sin(a1), cos(a1)], [3, 3])
    dy = reshape([cos(a2), 0.0d0, -sin(a2), 0.0d0, 1.0d0, 0.0d0, sin(a2), 0.0d0, cos(a2)], [3, 3])
    dz = reshape([cos(a3), sin(a3), 0.0d0, -sin(a3), cos(a2), 0.0d0, 0.0d0, 0.0d0, 1.0d0], [3, 3])
    do i = 1, 3
      do j = 1, 3
        acc = 0.0d0
        do k = 1, 3
          acc = acc + dx(i, k) * dy(k, j)
        end do
        t(i, j) = acc
      end do
    end do
    do i = 1, 3
      do j = 1, 3
        acc = 0.0d0
        do k = 1, 3
          acc = acc + t(i, k) * dz(k, j)
        end do
        dm(i, j) = acc
      end do
    end do
    do i = 1, 3
      do j = 1, 3
        acc = 0.0d0
        do k = 1, 3
          if (k == j) then
            acc = acc + dm(i, k) * v(j)
          else
            acc = acc + dm(i, k) * 0.0d0
          end if
        end do
        w(i, j) = acc
      end do
    end do
    do i = 1, 3
      do j = 1, 3
        acc = 0.0d0
        do k = 1, 3
          acc = acc + w(i, k) * dm(j, k)
        end do
        a(3 * (i - 1) + j) = acc * scale
      end do
    end do
  end subroutine aniso_tensor

  ! per-tet tensors (9, nT): conductivity, and the curl weight = entrywise
  ! reciprocal of the permeability tensor, zero entries kept (:621-633)
  subroutine tet_tensors(c, nT, tet_phy, sig9, nu9, ierr)
    type(cntl_t), intent(in) :: c
    integer, intent(in) :: nT
    integer(c_int), intent(in) :: tet_phy(nT)
    real(c_double), intent(out) :: sig9(9, nT), nu9(9, nT)
    integer, intent(out) :: ierr
    real(8) :: s(9, c%dom_num), mg(9), nu(9, c%dom_num)
    integer :: i, d, q
    logical :: found
    ierr = 0
    do d = 1, c%dom_num
      call aniso_tensor(c%adom(2:4, d), c%adom(5:7, d), 1.0d0, s(:, d))
      call aniso_tensor(c%adom(14:16, d), c%adom(17:19, d), mu0, mg)
      do q = 1, 9
        if (mg(q) == 0.0d0) then
          nu(q, d) = 0.0d0
        else
          nu(q, d) = 1.0d0 / mg(q)
        end if
      end do
    end do
    do i = 1, nT
      found = .false.
      do d = 1, c%dom_num
        if (tet_phy(i) == nint(c%adom(1, d))) then
          sig9(:, i) = s(:, d)
          nu9(:, i) = nu(:, d)
          found = .true.
          exit
        end if
      end do
      if (.not. found) then
        ierr = 1; return
      end if
    end do
  end subroutine tet_tensors

  ! ---- .field / .field2 ---------------------------------------------------
  subroutine fw_open(w, field_path, field2_path, nrec, o_f, gates, sig_parameter)
    type(field_writer_t), intent(out) :: w
    character(*), intent(in) :: field_path, field2_path
    integer, intent(in) :: nrec, o_f(9)
    real(8), intent(in) :: gates(:)
    real(8), intent(in), optional :: sig_parameter
    integer :: k
    character(16) :: a, b
    if (present(sig_parameter)) then
      w%rescale = .true.
      w%sp2 = sig_parameter**2.0d0
      w%spk = pi**log10(sig_parameter)
    end if
    w%nrec = nrec
    w%nsel = 0
    do k = 1, 9
      if (o_f(k) == 1) then
        w%nsel = w%nsel + 1
        w%o_v(w%nsel) = k
      end if
    end do
    write(a, '(i0)') nrec
    write(b, '(i0)') w%nsel
    w%fmt = '(i6,e15.5,2x,'//trim(a)//'('//trim(b)//'e15.5,2x))'
    allocate(w%gates(size(gates)), w%gate_ok(size(gates)), w%prev(9, nrec))
    w%gates = gates
    w%gate_ok = .false.
    w%nrow = 0
    open(newunit=w%u1, file=field_path, status='replace', action='write')
    open(newunit=w%u2, file=field2_path, status='replace', action='write')
  end subroutine fw_open

  ! one output row: vals(9, nrec) = ex ey ez dBx/dt dBy/dt dBz/dt bx by bz
  subroutine fw_row(w, step, t, vals)
    type(field_writer_t), intent(inout) :: w
    integer, intent(in) :: step
    real(8), intent(in) :: t, vals(9, w%nrec)
    real(8) :: g(9, w%nrec), a, b
    integer :: i, j, k
    w%nrow = w%nrow + 1
    write(w%u1, w%fmt) step, t, ((scaled(w, vals(w%o_v(k), j)), k = 1, w%nsel), j = 1, w%nrec)
    do i = 1, size(w%gates)
      if (t >= w%gates(i) .and. .not. w%gate_ok(i)) then
        if (w%nrow == 1) exit
        do j = 1, w%nrec
          do k = 1, 9
            a = (w%prev(k, j) - vals(k, j)) / (w%tprev - t)
            b = w%prev(k, j) - a * w%tprev
            g(k, j) = a * w%gates(i) + b
          end do
        end do
        write(w%u2, w%fmt) i, w%gates(i), ((scaled(w, g(w%o_v(k), j)), k = 1, w%nsel), j = 1, w%nrec)
        w%gate_ok(i) = .true.
      end if
    end do
    w%prev = vals
    w%tprev = t
  end subroutine fw_row

  pure real(8) function scaled(w, v)
    type(field_writer_t), intent(in) :: w
    real(8), intent(in) :: v
    if (w%rescale) then
      scaled = v / w%sp2 * w%spk
    else
      scaled = v
    end if
  end function scaled

  subroutine fw_close(w)
    type(field_writer_t), intent(inout) :: w
    if (w%u1 >= 0) close(w%u1)
    if (w%u2 >= 0) close(w%u2)
    w%u1 = -1
    w%u2 = -1
  end subroutine fw_close

  ! ---- per-tet field slices (msh_output, src/FETD_FWD.f90:1713-1864) --------
  ! Tetrahedra (element type 4) of a gmsh 4.x .msh and their tags in file
  ! order: the $Elements walk of msh_output (:1744-1788), each block's
  ! "dim tag type count" header, then the first number of every element line.
  subroutine msh_tet_count(path, ntet, tags, ierr, msg)
    character(*), intent(in) :: path
    integer, intent(out) :: ntet, ierr
    integer, allocatable, intent(out) :: tags(:)
    character(*), intent(out) :: msg
    integer :: u, ios, pass, nb, ib, d1, d2, et, cnt, j, tag
    character(4096) :: line
    ierr = 0; msg = ' '; ntet = 0
    open(newunit=u, file=path, status='old', action='read', iostat=ios)
    if (ios /= 0) then
      ierr = 1; msg = 'cannot open '//trim(path); return
    end if
    do pass = 1, 2
      rewind(u)
      do
        read(u, '(a)', iostat=ios) line
        if (ios /= 0) then
          ierr = 2; msg = 'cell number not found'; close(u); return
        end if
        if (line(1:9) == '$Elements') exit
      end do
      read(u, *) nb
      if (pass == 2) allocate(tags(ntet))
      ntet = 0
      do ib = 1, nb
        read(u, *) d1, d2, et, cnt
        do j = 1, cnt
          read(u, *) tag
          if (et == 4) then
            ntet = ntet + 1
            if (pass == 2) tags(ntet) = tag
          end if
        end do
      end do
    end do
    close(u)
  end subroutine msh_tet_count

  ! dst = a byte copy of src followed by one $ElementData section per
  ! selected component k (o_f(k) == 1): |f9(k, i)| per tetrahedron as
  ! "(i10,e16.8)" rows after the gmsh header lines (:1789-1862).  f9(9, nT):
  ! E, curl E, B in the mesh's tet order.
  subroutine write_msh_slice(src, dst, nT, f9, o_f, ierr, msg)
    character(*), intent(in) :: src, dst
    integer, intent(in) :: nT, o_f(9)
    real(8), intent(in) :: f9(9, nT)
    integer, intent(out) :: ierr
    character(*), intent(out) :: msg
    character(2), parameter :: names(9) = ['Ex', 'Ey', 'Ez', 'Dx', 'Dy', 'Dz', 'Bx', 'By', 'Bz']
    integer, allocatable :: tags(:)
    character(len=:), allocatable :: buf
    integer :: ntet, u, k, i, nbytes, ios
    call msh_tet_count(src, ntet, tags, ierr, msg)
    if (ierr /= 0) return
    if (ntet /= nT) then
      ierr = 3; write(msg, '(a,i0,a,i0)') 'the .msh has ', ntet, ' tetrahedra, the fields ', nT; return
    end if
    inquire(file=src, size=nbytes)
    allocate(character(len=nbytes) :: buf)
    open(newunit=u, file=src, access='stream', form='unformatted', status='old', action='read', iostat=ios)
    if (ios == 0) read(u, iostat=ios) buf
    close(u)
    if (ios /= 0) then
      ierr = 1; msg = 'cannot read '//trim(src); return
    end if
    open(newunit=u, file=dst, access='stream', form='unformatted', status='replace', action='write', iostat=ios)
    if (ios /= 0) then
      ierr = 1; msg = 'cannot write '//trim(dst); return
    end if
    write(u) buf
    close(u)
    open(newunit=u, file=dst, position='append', action='write', form='formatted')
    do k = 1, 9
      if (o_f(k) /= 1) cycle
      write(u, '(a)') '$ElementData'
      write(u, '(a)') '1'
      write(u, '(a)') '"'//names(k)//' data"'
      write(u, '(a)') '1'
      write(u, '(a)') '0.0'
      write(u, '(a)') '3'
      write(u, '(a)') '0'
      write(u, '(a)') '1'
      write(u, '(i12)') nT
      do i = 1, nT
        write(u, '(i10,e16.8)') tags(i), abs(f9(k, i))
      end do
      write(u, '(a)') '$EndElementData'
    end do
    close(u)
  end subroutine write_msh_slice
end module fetd_io
