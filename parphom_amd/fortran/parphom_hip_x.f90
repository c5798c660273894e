! parphom_hip.x - the parphom.x host program on the MI355X library.
!
!   parphom_hip.x inp.dat
!
! Follows main.F90:40-70: read_input (read_input.F90:35-333 and the
! sanitize_input rules of :415-516), read_lammps_data (lammps_read.F90:66-192,
! positional), read_q_file (q_read.F90:82-93), read_force_constants (one
! contiguous HDF5 read of the (N,N,3,3) dataset, force_read.F90:93-158), then
! the q loop of calculations.F90:49-53 as chunks of pph_run calls, writing one
! HDF5 group per q in the layout of write_output.F90:95-464 (evec only with
! `compute eigvecs : yes`, vel only with `group velocity : yes`).  Two keys
! parphom.x does not know are accepted: `precision : S|D` and `batch : B`.
program parphom_hip_x
  use iso_c_binding
  use parphom_hip
  implicit none

  type :: input_t
    character(len=1024) :: lammps_location = './', lammps_name = 'lammps.dat'
    character(len=64)   :: atom_style = 'A'
    integer(c_int64_t)  :: atom_types = 1, natom = 1
    character(len=1024) :: q_location = './', q_name = 'q_file.dat'
    integer(c_int64_t)  :: nqpt = 1
    logical             :: evec_comp = .false., comp_vel = .false., print_progress = .true.
    character(len=1)    :: vel_method = 'A', jobz = 'N', range = 'A', precision = 'D'
    integer(c_int64_t)  :: il = 1, iu = 1, batch = 0
    real(c_double)      :: vl = 0.0d0, vu = 0.0d0
    real(c_double)      :: abstol = real(1.0e-15, c_double), orfac = real(1.0e-15, c_double)
    character(len=1024) :: fc_location = './', fc_name = 'FC_hdf5', fc_dataset = 'force_constants'
    character(len=1024) :: output_name = 'results', output_location = './'
  end type

  type(input_t) :: inp
  character(len=1024) :: inp_file, out_path
  real(c_double), allocatable, target :: lat(:, :), pos(:, :), mass(:), q(:, :), fc(:, :, :, :)
  integer(c_int64_t), allocatable, target :: types(:)
  real(c_double), allocatable, target :: eval(:, :), vel(:, :, :), qchunk(:, :)
  complex(c_double_complex), allocatable, target :: evec(:, :, :)
  integer(c_int64_t), allocatable, target :: m(:), nz(:)
  integer(c_int32_t), allocatable, target :: info(:)
  type(pph_system) :: sys
  type(pph_opts) :: opts
  type(pph_out) :: out
  type(c_ptr) :: h
  integer(c_int64_t) :: n, chunk, q0, nb, k, fid
  integer(c_int) :: ierr
  type(c_ptr) :: pe, pv

  if (command_argument_count() < 1) then
    write(*, '(a)') 'usage: parphom_hip.x inp.dat'
    stop 2
  end if
  call get_command_argument(1, inp_file)
  call read_input(trim(inp_file), inp)
  call sanitize_input(inp)
  call read_lammps_data(trim(inp%lammps_location) // trim(inp%lammps_name), inp, lat, pos, mass, types)
  call read_q_file(trim(inp%q_location) // trim(inp%q_name), inp%nqpt, q)
  n = 3 * inp%natom
  allocate(fc(3, 3, inp%natom, inp%natom))
  ierr = pph_h5_read_fc(trim(inp%fc_location) // trim(inp%fc_name) // c_null_char, &
                        trim(inp%fc_dataset) // c_null_char, inp%natom, fc)
  if (ierr /= 0) call die('Error reading force constants: ' // c_to_string(pph_h5_last_error()))
  if (inp%print_progress) write(*, '(a,i0,a,i0,a,a,a,a)') 'Inputs read: natom=', inp%natom, &
      ' nqpt=', inp%nqpt, ' range=', inp%range, ' precision=', inp%precision

  sys%natom = inp%natom
  sys%ntypes = inp%atom_types
  sys%lat = c_loc(lat)
  sys%pos = c_loc(pos)
  sys%mass = c_loc(mass)
  sys%type = c_loc(types)
  opts%jobz = inp%jobz
  opts%range = inp%range
  opts%vel_method = inp%vel_method
  opts%precision = inp%precision
  opts%vel = merge(1, 0, inp%comp_vel)
  opts%device = 0
  opts%vl = inp%vl
  opts%vu = inp%vu
  opts%il = inp%il
  opts%iu = inp%iu
  opts%abstol = inp%abstol
  opts%orfac = inp%orfac
  opts%batch = inp%batch
  opts%streams = 0
  ierr = pph_create(sys, fc, opts, h)
  if (ierr /= PPH_OK) call die('pph_create: ' // c_to_string(pph_last_error(c_null_ptr)))
  deallocate(fc)

  out_path = trim(inp%output_location) // trim(inp%output_name)
  ierr = pph_h5_open_output(trim(out_path) // c_null_char, 1_c_int32_t, fid)
  if (ierr /= 0) call die('cannot create ' // trim(out_path))

  ! the q-points in chunks of the library's batch (host buffers for one chunk)
  chunk = merge(inp%batch, 32_c_int64_t, inp%batch > 0)
  chunk = min(chunk, inp%nqpt)
  allocate(eval(n, chunk), m(chunk), nz(chunk), info(chunk), qchunk(chunk, 3))
  if (inp%jobz == 'V') allocate(evec(n, n, chunk))
  if (inp%comp_vel) allocate(vel(3, n, chunk))
  do q0 = 1, inp%nqpt, chunk
    nb = min(chunk, inp%nqpt - q0 + 1)
    qchunk(1:nb, :) = q(q0:q0 + nb - 1, :)
    out%eval_cm = c_loc(eval)
    out%ld_eval = n
    out%m = c_loc(m)
    out%nz = c_loc(nz)
    out%info = c_loc(info)
    out%evec = c_null_ptr
    out%vel = c_null_ptr
    out%evec_cols = n
    if (allocated(evec)) out%evec = c_loc(evec)
    if (allocated(vel)) out%vel = c_loc(vel)
    ierr = pph_run(h, nb, qchunk, chunk, out)
    if (ierr /= PPH_OK) call die('pph_run: ' // c_to_string(pph_last_error(h)))
    do k = 1, nb
      pe = c_null_ptr
      pv = c_null_ptr
      if (inp%evec_comp .and. allocated(evec)) pe = c_loc(evec(1, 1, k))
      if (inp%comp_vel .and. allocated(vel)) pv = c_loc(vel(1, 1, k))
      ierr = pph_h5_write_group(fid, q0 + k - 1, q(q0 + k - 1, :), eval(:, k), m(k), pe, n, &
                                nz(k), pv)
      if (ierr /= 0) call die('writing the output group: ' // c_to_string(pph_h5_last_error()))
      if (info(k) /= 0) write(*, '(a,i0,a,i0)') 'q-point ', q0 + k - 1, ': eigensolver info ', info(k)
    end do
    if (inp%print_progress) write(*, '(a,i0,a,i0,a,i0,a)') 'q-points ', q0, '..', q0 + nb - 1, &
        ' of ', inp%nqpt, ' done'
  end do
  ierr = pph_h5_close(fid)
  ierr = pph_destroy(h)
  write(*, '(a,a)') 'output written to ', trim(out_path)

contains

  subroutine die(msg)
    character(len=*), intent(in) :: msg
    write(*, '(a)') msg
    error stop 1
  end subroutine

  pure function upper(s) result(u)
    character(len=*), intent(in) :: s
    character(len=len(s)) :: u
    integer :: i, c
    u = s
    do i = 1, len(s)
      c = iachar(s(i:i))
      if (c >= iachar('a') .and. c <= iachar('z')) u(i:i) = achar(c - 32)
    end do
  end function

  ! the key with spaces and underscores equal, as read_input.F90 lists both spellings
  pure logical function key_is(key, name)
    character(len=*), intent(in) :: key, name
    character(len=len(key)) :: a
    character(len=len(name)) :: b
    integer :: i
    a = key
    b = name
    do i = 1, len(a)
      if (a(i:i) == '_') a(i:i) = ' '
    end do
    do i = 1, len(b)
      if (b(i:i) == '_') b(i:i) = ' '
    end do
    key_is = trim(a) == trim(b)
  end function

  logical function yes_no(v, default, what)
    character(len=*), intent(in) :: v, what
    logical, intent(in) :: default
    select case (trim(v))
    case ('yes', 'true', 'Yes', 'True', 'YES', 'TRUE')
      yes_no = .true.
    case ('no', 'false', 'No', 'False', 'NO', 'FALSE')
      yes_no = .false.
    case default
      write(*, '(a,a)') 'Could not interpret command ', trim(v)
      write(*, '(a)') what
      yes_no = default
    end select
  end function

  integer(c_int64_t) function to_int(v)
    character(len=*), intent(in) :: v
    integer :: ios
    read(v, *, iostat=ios) to_int
    if (ios /= 0) call die('cannot read an integer from "' // trim(v) // '"')
  end function

  real(c_double) function to_real(v)
    character(len=*), intent(in) :: v
    integer :: ios
    read(v, *, iostat=ios) to_real
    if (ios /= 0) call die('cannot read a real from "' // trim(v) // '"')
  end function

  ! read_input.F90:35-333: one KEY : value per line, '#' / '!' comments
  subroutine read_input(path, inp)
    character(len=*), intent(in) :: path
    type(input_t), intent(inout) :: inp
    character(len=2048) :: line
    character(len=1024) :: key, v
    integer :: u, ios, p
    open(newunit=u, file=path, status='old', action='read', iostat=ios)
    if (ios /= 0) call die('Error reading input file ' // path)
    do
      read(u, '(a)', iostat=ios) line
      if (ios /= 0) exit
      line = adjustl(line)
      if (len_trim(line) == 0) cycle
      if (line(1:1) == '#' .or. line(1:1) == '!') cycle
      p = index(line, '!')
      if (p > 0) line = line(1:p - 1)
      p = index(line, '#')
      if (p > 0) line = line(1:p - 1)
      p = index(line, ':')
      if (p > 0) then
        key = upper(line(1:p - 1))
        v = adjustl(line(p + 1:))
      else
        key = ''
        v = adjustl(line)
      end if
      if (key_is(key, 'LAMMPS FILE LOCATION')) then
        inp%lammps_location = v
      else if (key_is(key, 'LAMMPS FILE NAME')) then
        inp%lammps_name = v
      else if (key_is(key, 'NATOM')) then
        inp%natom = to_int(v)
      else if (key_is(key, 'ATOM STYLE')) then
        inp%atom_style = v
      else if (key_is(key, 'ATOM TYPES')) then
        inp%atom_types = to_int(v)
      else if (key_is(key, 'FORCE CONSTANT FILE LOCATION')) then
        inp%fc_location = v
      else if (key_is(key, 'FORCE CONSTANT FILE NAME')) then
        inp%fc_name = v
      else if (key_is(key, 'FORCE CONSTANT DATASET')) then
        inp%fc_dataset = v
      else if (key_is(key, 'Q FILE LOCATION')) then
        inp%q_location = v
      else if (key_is(key, 'Q FILE NAME')) then
        inp%q_name = v
      else if (key_is(key, 'NQPT')) then
        inp%nqpt = to_int(v)
      else if (key_is(key, 'COMPUTE EIGVECS')) then
        inp%evec_comp = yes_no(v, .false., 'Eigenvectors will not be computed.')
        inp%jobz = merge('V', 'N', inp%evec_comp)
      else if (key_is(key, 'GROUP VELOCITY')) then
        inp%comp_vel = yes_no(v, .false., 'Velocity will not be computed.')
      else if (key_is(key, 'VELOCITY METHOD')) then
        select case (trim(v))
        case ('cd', 'CD', 'central_difference', 'CENTRAL_DIFFERENCE')
          inp%vel_method = 'C'
        case ('A', 'a', 'Analytic', 'analytic')
          inp%vel_method = 'A'
        case default
          write(*, '(a,a)') 'Could not interpret command ', trim(v)
          write(*, '(a)') 'Velocity will be computed using analytic derivative.'
          inp%vel_method = 'A'
        end select
      else if (key_is(key, 'RANGE')) then
        if (trim(v) == 'A' .or. trim(v) == 'V' .or. trim(v) == 'I') inp%range = v(1:1)
      else if (key_is(key, 'MAX EIGVAL')) then
        inp%vu = to_real(v)
      else if (key_is(key, 'MIN EIGVAL')) then
        inp%vl = to_real(v)
      else if (key_is(key, 'MAX INDEX')) then
        inp%iu = to_int(v)
      else if (key_is(key, 'MIN INDEX')) then
        inp%il = to_int(v)
      else if (key_is(key, 'ABSTOL')) then
        inp%abstol = to_real(v)
      else if (key_is(key, 'ORFAC')) then
        inp%orfac = to_real(v)
      else if (key_is(key, 'MB') .or. key_is(key, 'NB')) then
        continue                 ! block-cyclic blocking: no meaning on one GPU
      else if (key_is(key, 'OUTPUT FILE NAME')) then
        inp%output_name = v
      else if (key_is(key, 'OUTPUT FILE LOCATION')) then
        inp%output_location = v
      else if (key_is(key, 'PRINT PROGRESS')) then
        inp%print_progress = yes_no(v, .true., 'Progress will be printed.')
      else if (key_is(key, 'PRECISION')) then
        select case (trim(upper(v)))
        case ('S', 'FP32', 'SINGLE', 'MIXED')
          inp%precision = 'S'
        case default
          inp%precision = 'D'
        end select
      else if (key_is(key, 'BATCH')) then
        inp%batch = to_int(v)
      else
        write(*, '(a,a)') 'Command ', trim(key) // ' not recognised'
        write(*, '(a)') 'Ignoring this command'
      end if
    end do
    close(u)
  end subroutine

  ! sanitize_input, read_input.F90:415-516 (one process)
  subroutine sanitize_input(inp)
    type(input_t), intent(inout) :: inp
    character(len=1024) :: name
    logical :: there
    integer :: k
    call slash(inp%lammps_location)
    call slash(inp%q_location)
    call slash(inp%output_location)
    call slash(inp%fc_location)
    select case (inp%range)
    case ('V')
      inp%il = 1
      inp%iu = inp%natom
    case ('I')
      inp%vl = 0
      inp%vu = 0
    case default
      inp%il = 1
      inp%iu = inp%natom
      inp%vl = 0
      inp%vu = 0
    end select
    if (inp%range == 'I' .and. inp%il > inp%iu) &
      call die('Invalid input for eigenvalue indices to be computed')
    if (inp%range == 'V' .and. inp%vl > inp%vu) &
      call die('Invalid input for the interval in which eigenvalue is to be computed')
    if (inp%comp_vel) inp%jobz = 'V'      ! vectors computed, written only with evec_comp
    inquire(file=trim(inp%output_location) // trim(inp%output_name) // '.hdf5', exist=there)
    if (there) then
      write(*, '(a)') 'File ' // trim(inp%output_name) // '.hdf5 already exists at ' // &
          trim(inp%output_location)
      k = 0
      do
        k = k + 1
        write(name, '(a,a,i0)') trim(inp%output_name), '_v', k
        inquire(file=trim(inp%output_location) // trim(name) // '.hdf5', exist=there)
        if (.not. there) exit
      end do
      inp%output_name = trim(name) // '.hdf5'
      write(*, '(a)') 'Output file being renamed as ' // trim(inp%output_name)
    else
      inp%output_name = trim(inp%output_name) // '.hdf5'
    end if
  end subroutine

  subroutine slash(loc)
    character(len=*), intent(inout) :: loc
    integer :: l
    l = len_trim(loc)
    if (l == 0) then
      loc = '/'
    else if (loc(l:l) /= '/') then
      loc = loc(1:l) // '/'
    end if
  end subroutine

  subroutine skip(u, k)
    integer, intent(in) :: u, k
    integer :: i, ios
    character(len=16) :: dummy
    do i = 1, k
      read(u, '(a)', iostat=ios) dummy
      if (ios /= 0) call die('unexpected end of LAMMPS file')
    end do
  end subroutine

  ! lammps_read.F90:66-192: skip 5 lines; xlo xhi, ylo yhi, zlo zhi, xy xz yz;
  ! skip 3; the masses; skip 3; one line per atom in the given style
  subroutine read_lammps_data(path, inp, lat, pos, mass, types)
    character(len=*), intent(in) :: path
    type(input_t), intent(in) :: inp
    real(c_double), allocatable, intent(out) :: lat(:, :), pos(:, :), mass(:)
    integer(c_int64_t), allocatable, intent(out) :: types(:)
    real(c_double) :: x1, x2, y1, y2, z1, z2, xy, xz, yz, x(3)
    integer(c_int64_t) :: a, t, id, layer, img(3)
    real(c_double) :: q_dummy
    integer :: u, ios
    open(newunit=u, file=path, status='old', action='read', iostat=ios)
    if (ios /= 0) call die('Error reading LAMMPS file ' // path)
    call skip(u, 5)
    read(u, *, iostat=ios) x1, x2
    if (ios == 0) read(u, *, iostat=ios) y1, y2
    if (ios == 0) read(u, *, iostat=ios) z1, z2
    if (ios == 0) read(u, *, iostat=ios) xy, xz, yz
    if (ios /= 0) call die('malformed LAMMPS box lines')
    allocate(lat(3, 3), mass(inp%atom_types), pos(inp%natom, 3), types(inp%natom))
    ! lat(i, :) = lattice vector i
    lat(1, :) = [x2 - x1, 0.0d0, 0.0d0]
    lat(2, :) = [xy, y2 - y1, 0.0d0]
    lat(3, :) = [xz, yz, z2 - z1]
    call skip(u, 3)
    do t = 1, inp%atom_types
      read(u, *, iostat=ios) id, mass(t)
      if (ios /= 0) call die('malformed LAMMPS mass line')
    end do
    call skip(u, 3)
    do a = 1, inp%natom
      select case (trim(inp%atom_style))
      case ('atomic', 'Atomic', 'ATOMIC')
        read(u, *, iostat=ios) id, types(a), x, img
      case ('molecular', 'Molecular', 'MOLECULAR')
        read(u, *, iostat=ios) id, layer, types(a), x, img
      case ('full', 'Full', 'FULL')
        read(u, *, iostat=ios) id, layer, types(a), q_dummy, x, img
      case default
        call die('Atom style not supported. Supported styles are atomic and molecular')
      end select
      if (ios /= 0) call die('malformed LAMMPS atom line')
      pos(a, :) = x
    end do
    close(u)
  end subroutine

  ! q_read.F90:82-93: nqpt list-directed rows of >= 3 reals
  subroutine read_q_file(path, nqpt, q)
    character(len=*), intent(in) :: path
    integer(c_int64_t), intent(in) :: nqpt
    real(c_double), allocatable, intent(out) :: q(:, :)
    integer :: u, ios
    integer(c_int64_t) :: i
    open(newunit=u, file=path, status='old', action='read', iostat=ios)
    if (ios /= 0) call die('Error reading q-point file ' // path)
    allocate(q(nqpt, 3))
    do i = 1, nqpt
      read(u, *, iostat=ios) q(i, 1:3)
      if (ios /= 0) call die('q file has fewer rows than nqpt')
    end do
    close(u)
  end subroutine

end program parphom_hip_x
