! tlq_hydro.f90 -- a Fortran driver of the GPU time step, in the shape of the
! reference's main program (source/mainHD.f90:20-110): set up U on the host,
! fill the borders and primitives, then loop while t < tfin, writing an output
! whenever t >= tout, and a last one after the loop.  The time step itself runs
! on the GPU through libtlq (module tlq_c): either the fused step
! (tlq_advance) or the reference's thirteen calls one by one on the device
! arrays (mainHD.f90:74-99; per_routine = .true.).
!
! Settings come from a namelist file given as the first argument:
!   &mesh     nx, ny, nz, xl, xr, yl, yr, zl, zr
!   &physics  gamma, mu0, cfl, eta, cooling_limit, limiter, solver, precision,
!             cooling, bc(6), softening, tfin, cooling_file, self_gravity,
!             mg_levels, gconst, mg_tol, phifile (self-gravity: PHI starts from
!             phifile, a little-endian real(8) stream of PHI(0:nx+1,0:ny+1,
!             0:nz+1) whose ghost faces are the boundary values, else from 0;
!             TLQ_SG_MULTIPOLE sets the faces itself)
!   &run      problem, num_out, outfile, outputpath, max_steps, per_routine,
!             initfile, dry_run, nwinds, winds(:), nmasses, masses(4,:)
! problem: 'sod'    diagonal Sod tube of the reference's test harness
!                   (rho,p = 4,1 where x+y+z < 0, else 1,0.1795, at rest)
!          'sod_x'  planar Sod tube along x (rho,p = 1,1 for x < (xl+xr)/2,
!                   else 0.125,0.1)
!          'sedov'  point explosion of the harness: E0 = 1e5 in r < 0.15,
!                   rho = 1, p = 1.5e-20 (gamma-1) outside
!          'blast'  point explosion: E = 1 spread over the cells with
!                   r < 3 dx, rho = 1, p = 1e-5 outside
!          'file'   U read from initfile (little-endian real(8) stream of
!                   U(5,-1:nx+2,-1:ny+2,-1:nz+2))
!          'restart' continue a run from one of its outputs, restart_file:
!                   a BIN dump (UPrim with ghosts; the ghosts are kept, so
!                   the run continues to rounding) or a VTK file (interior
!                   cells; the ghosts come from the boundary conditions, as
!                   at a fresh start).  U is rebuilt from the primitives;
!                   t, it and the output number come from the dump's sidecar
!                   <restart_file>.time unless restart_t, restart_it,
!                   restart_nout are given.  With self-gravity, a BIN restart
!                   also reads PHIT from 'phi' + restart_file when that file
!                   exists and starts PHI from PHIT - PHIP.
!                   This is what source/warmstart.f90:1-10 set out to do
!                   (read a VTK output back), which the reference never
!                   finished.
! outfile: 'vtk', 'bin' or 'dat', written as source/HydroCore.f90:253-395
! does (file names EPV000Nnnn.vtk, points000Nnnn.bin / .dat), so the
! reference's analysis scripts read them unchanged.  VTK and BIN always hold
! real(8) values under the 'double' header.
! Every output also writes <file>.time, one line "t it nout", for restarts.
! With self-gravity the outputs also carry PHIT = PHIP + PHI as the -DGRAV
! build writes it (HydroCore.f90:284-290, 295-301, 360-372): BIN to 'phi'
! prefixed to the BIN path (e.g. phi./DATA/points000N001.bin; that folder
! must exist, as for the reference), a ninth DAT column, and a "SCALARS
! Potential" block in the VTK.
! dry_run = .true. writes the initial state's output without touching the GPU.
module driver_setup
  use iso_c_binding
  use tlq_c
  implicit none

  real(8), parameter :: AMU = 1.660538782d-24, KB = 1.380650400d-16, PI = 3.14159265358979d0

  integer, parameter :: MAXW = 64, MAXM = 64

  ! &mesh
  integer :: nx = 32, ny = 32, nz = 32
  real(8) :: xl = -1d0, xr = 1d0, yl = -1d0, yr = 1d0, zl = -1d0, zr = 1d0
  ! &physics
  real(8) :: gamma = 5d0/3d0, mu0 = 1d0, cfl = 0.4d0, eta = 0.5d-2, cooling_limit = 0.5d0
  real(8) :: softening = 0d0, tfin = 1d0
  integer :: limiter = TLQ_LIM_VANLEER, solver = TLQ_SOLVER_HLLC, precision = TLQ_FP64
  integer :: cooling = TLQ_COOL_OFF
  integer :: self_gravity = TLQ_SG_OFF, mg_levels = 0
  real(8) :: gconst = 6.67430d-8, mg_tol = 0d0
  integer :: bc(6) = TLQ_BC_REFLECTIVE
  character(len=512) :: cooling_file = '', phifile = ''
  ! &run
  character(len=32) :: problem = 'sod'
  integer :: num_out = 1, max_steps = 0, nwinds = 0, nmasses = 0
  character(len=8) :: outfile = 'vtk'
  character(len=512) :: outputpath = './DATA/', initfile = '', restart_file = ''
  real(8) :: restart_t = -1d0
  integer :: restart_it = -1, restart_nout = -1
  logical :: restart_fill = .true.   ! the dump has no ghosts (VTK)
  logical :: per_routine = .false., dry_run = .false.
  type(tlq_wind_t) :: winds(MAXW)
  real(8) :: masses(4, MAXM) = 0d0

  namelist /mesh/ nx, ny, nz, xl, xr, yl, yr, zl, zr
  namelist /physics/ gamma, mu0, cfl, eta, cooling_limit, limiter, solver, precision, cooling, &
       bc, softening, tfin, cooling_file, self_gravity, mg_levels, gconst, phifile, mg_tol
  namelist /run/ problem, num_out, outfile, outputpath, max_steps, per_routine, initfile, &
       dry_run, nwinds, winds, nmasses, masses, restart_file, restart_t, restart_it, restart_nout

  real(8) :: dx, dy, dz

contains

  subroutine read_settings(path)
    character(*), intent(in) :: path
    integer :: u, ios
    open(newunit=u, file=path, status='old', action='read', iostat=ios)
    if (ios /= 0) then
       write(*, '(a,a)') 'cannot open settings file ', trim(path)
       error stop 2
    end if
    read(u, nml=mesh, iostat=ios)
    if (ios /= 0) then
       write(*, '(a)') 'bad &mesh namelist'
       error stop 2
    end if
    rewind(u)
    read(u, nml=physics, iostat=ios)
    if (ios /= 0) then
       write(*, '(a)') 'bad &physics namelist'
       error stop 2
    end if
    rewind(u)
    read(u, nml=run, iostat=ios)
    if (ios /= 0) then
       write(*, '(a)') 'bad &run namelist'
       error stop 2
    end if
    close(u)
    dx = (xr - xl) / nx
    dy = (yr - yl) / ny
    dz = (zr - zl) / nz
  end subroutine read_settings

  function make_params() result(p)
    type(tlq_params_t) :: p
    p%nx = nx
    p%ny = ny
    p%nz = nz
    p%xl = xl
    p%xr = xr
    p%yl = yl
    p%yr = yr
    p%zl = zl
    p%zr = zr
    p%gamma = gamma
    p%mu0 = mu0
    p%cfl = cfl
    p%eta = eta
    p%cooling_limit = cooling_limit
    p%limiter = limiter
    p%bc = bc
    p%solver = solver
    p%precision = precision
    p%cooling = cooling
    p%softening = softening
    p%self_gravity = self_gravity
    p%mg_levels = mg_levels
    p%gconst = gconst
    p%mg_tol = mg_tol
  end function make_params

  ! cell centres, as xcoord/ycoord/zcoord (source/HydroCore.f90:35-75, one rank)
  real(8) function xc(i)
    integer, intent(in) :: i
    xc = xl + dx * i - dx / 2d0
  end function xc
  real(8) function yc(j)
    integer, intent(in) :: j
    yc = yl + dy * j - dy / 2d0
  end function yc
  real(8) function zc(k)
    integer, intent(in) :: k
    zc = zl + dz * k - dz / 2d0
  end function zc

  subroutine set_cell(U, i, j, k, rho, vx, vy, vz, p)
    real(8), intent(inout) :: U(5, -1:nx+2, -1:ny+2, -1:nz+2)
    integer, intent(in) :: i, j, k
    real(8), intent(in) :: rho, vx, vy, vz, p
    U(1, i, j, k) = rho
    U(2, i, j, k) = rho * vx
    U(3, i, j, k) = rho * vy
    U(4, i, j, k) = rho * vz
    U(5, i, j, k) = 0.5d0 * rho * (vx**2 + vy**2 + vz**2) + p / (gamma - 1d0)
  end subroutine set_cell

  ! Initial conditions over every cell, ghosts included (initflow).
  subroutine initflow(U)
    real(8), intent(inout) :: U(5, -1:nx+2, -1:ny+2, -1:nz+2)
    integer :: i, j, k, ncells, u_in
    real(8) :: r, pin, xmid
    select case (trim(problem))
    case ('sod')
       do k = -1, nz + 2
          do j = -1, ny + 2
             do i = -1, nx + 2
                if (xc(i) + yc(j) + zc(k) < 0d0) then
                   call set_cell(U, i, j, k, 4d0, 0d0, 0d0, 0d0, 1d0)
                else
                   call set_cell(U, i, j, k, 1d0, 0d0, 0d0, 0d0, 0.1795d0)
                end if
             end do
          end do
       end do
    case ('sod_x')
       xmid = 0.5d0 * (xl + xr)
       do k = -1, nz + 2
          do j = -1, ny + 2
             do i = -1, nx + 2
                if (xc(i) < xmid) then
                   call set_cell(U, i, j, k, 1d0, 0d0, 0d0, 0d0, 1d0)
                else
                   call set_cell(U, i, j, k, 0.125d0, 0d0, 0d0, 0d0, 0.1d0)
                end if
             end do
          end do
       end do
    case ('sedov')
       pin = (1d5 / ((4d0 / 3d0) * PI * 0.15d0**3)) * (gamma - 1d0)
       do k = -1, nz + 2
          do j = -1, ny + 2
             do i = -1, nx + 2
                r = sqrt(xc(i)**2 + yc(j)**2 + zc(k)**2)
                if (r < 0.15d0) then
                   call set_cell(U, i, j, k, 1d0, 0d0, 0d0, 0d0, pin)
                else
                   call set_cell(U, i, j, k, 1d0, 0d0, 0d0, 0d0, 1.5d-20 * (gamma - 1d0))
                end if
             end do
          end do
       end do
    case ('blast')
       ncells = 0
       do k = 1, nz
          do j = 1, ny
             do i = 1, nx
                if (sqrt(xc(i)**2 + yc(j)**2 + zc(k)**2) < 3d0 * dx) ncells = ncells + 1
             end do
          end do
       end do
       pin = (gamma - 1d0) * 1d0 / (ncells * dx * dy * dz)
       do k = -1, nz + 2
          do j = -1, ny + 2
             do i = -1, nx + 2
                r = sqrt(xc(i)**2 + yc(j)**2 + zc(k)**2)
                if (r < 3d0 * dx .and. i >= 1 .and. i <= nx .and. j >= 1 .and. j <= ny .and. &
                     k >= 1 .and. k <= nz) then
                   call set_cell(U, i, j, k, 1d0, 0d0, 0d0, 0d0, pin)
                else
                   call set_cell(U, i, j, k, 1d0, 0d0, 0d0, 0d0, 1d-5)
                end if
             end do
          end do
       end do
    case ('file')
       open(newunit=u_in, file=trim(initfile), status='old', access='stream', &
            form='unformatted', convert='LITTLE_ENDIAN', action='read')
       read(u_in) U
       close(u_in)
    case ('restart')
       call read_restart(U)
    case default
       write(*, '(a,a)') 'unknown problem ', trim(problem)
       error stop 2
    end select
  end subroutine initflow

  ! U from the primitives of a dump (the inverse of calc_prims,
  ! source/HydroCore.f90:14-21).  A VTK dump holds the interior cells only:
  ! its ghosts start as copies of the nearest interior cell and the
  ! boundary fill at start-up (tlq_init_state) sets them.
  subroutine read_restart(U)
    real(8), intent(inout) :: U(5, -1:nx+2, -1:ny+2, -1:nz+2)
    real(8), allocatable :: W(:, :, :, :)
    integer :: u_in, n, i, j, k
    allocate(W(5, -1:nx+2, -1:ny+2, -1:nz+2))
    n = len_trim(restart_file)
    restart_fill = n > 4 .and. restart_file(max(n-3, 1):n) == '.vtk'
    if (restart_fill) then
       call read_vtk_prims(trim(restart_file), W)
       do k = -1, nz + 2
          do j = -1, ny + 2
             do i = -1, nx + 2
                W(:, i, j, k) = W(:, min(max(i, 1), nx), min(max(j, 1), ny), min(max(k, 1), nz))
             end do
          end do
       end do
    else
       open(newunit=u_in, file=trim(restart_file), status='old', access='stream', &
            form='unformatted', convert='LITTLE_ENDIAN', action='read')
       read(u_in) W
       close(u_in)
    end if
    do k = -1, nz + 2
       do j = -1, ny + 2
          do i = -1, nx + 2
             call set_cell(U, i, j, k, W(1, i, j, k), W(2, i, j, k), W(3, i, j, k), &
                  W(4, i, j, k), W(5, i, j, k))
          end do
       end do
    end do
  end subroutine read_restart

  ! The legacy binary VTK this driver writes (DIMENSIONS nx ny nz, then the
  ! interior Density, Pressure and Velocity as big-endian real(8)).
  subroutine read_vtk_prims(path, W)
    character(*), intent(in) :: path
    real(8), intent(inout) :: W(5, -1:nx+2, -1:ny+2, -1:nz+2)
    character(len=200) :: line
    integer :: u_in, i, j, k, dims(3)
    character(len=1) :: c
    open(newunit=u_in, file=path, status='old', access='stream', form='unformatted', &
         convert='BIG_ENDIAN', action='read')
    dims = 0
    do
       call next_line()
       if (line(1:10) == 'DIMENSIONS') then
          read(line(11:), *) dims
          if (any(dims /= [nx, ny, nz])) then
             write(*, '(a,3i6)') 'restart file has DIMENSIONS ', dims
             error stop 2
          end if
       else if (line(1:15) == 'SCALARS Density') then
          call next_line()
          read(u_in) (((W(1, i, j, k), i = 1, nx), j = 1, ny), k = 1, nz)
       else if (line(1:16) == 'SCALARS Pressure') then
          call next_line()
          read(u_in) (((W(5, i, j, k), i = 1, nx), j = 1, ny), k = 1, nz)
       else if (line(1:16) == 'VECTORS Velocity') then
          read(u_in) (((W(2, i, j, k), W(3, i, j, k), W(4, i, j, k), i = 1, nx), j = 1, ny), &
               k = 1, nz)
          exit
       end if
    end do
    close(u_in)
    if (dims(1) == 0) then
       write(*, '(a)') 'restart file has no DIMENSIONS line'
       error stop 2
    end if
  contains
    subroutine next_line()
      integer :: m, ios
      line = ''
      m = 0
      do
         read(u_in, iostat=ios) c
         if (ios /= 0) then
            write(*, '(a,a)') 'restart file ends early: ', path
            error stop 2
         end if
         if (c == char(10)) then
            if (m > 0) exit
            cycle
         end if
         m = m + 1
         if (m <= len(line)) line(m:m) = c
      end do
    end subroutine next_line
  end subroutine read_vtk_prims

  ! t, it and the output number of a restart: the namelist's values where
  ! given, else the dump's sidecar.
  subroutine restart_clock(t, it, nout)
    real(8), intent(out) :: t
    integer(c_int), intent(out) :: it
    integer, intent(out) :: nout
    integer :: u_in, ios, it_f, nout_f
    real(8) :: t_f
    t_f = 0d0
    it_f = 0
    nout_f = 0
    open(newunit=u_in, file=trim(restart_file)//'.time', status='old', action='read', iostat=ios)
    if (ios == 0) then
       read(u_in, *, iostat=ios) t_f, it_f, nout_f
       close(u_in)
       if (ios /= 0) then
          write(*, '(a)') 'bad restart sidecar ', trim(restart_file)//'.time'
          error stop 2
       end if
    else if (restart_t < 0d0) then
       write(*, '(a,a)') 'no sidecar and no restart_t for ', trim(restart_file)
       error stop 2
    end if
    t = t_f
    it = it_f
    nout = nout_f
    if (restart_t >= 0d0) t = restart_t
    if (restart_it >= 0) it = restart_it
    if (restart_nout >= 0) nout = restart_nout
  end subroutine restart_clock

  ! PHIT = PHIP + PHI (HLLC.f90:22): the point masses' potential
  ! (pointsource_gravity.f90:30-35, with the softening of tlq_params) added to
  ! the self-gravity potential PHI, on every cell of PHI(0:nx+1, ...).
  subroutine make_phit(PHI, PHIT)
    real(8), intent(in) :: PHI(0:nx+1, 0:ny+1, 0:nz+1)
    real(8), intent(out) :: PHIT(0:nx+1, 0:ny+1, 0:nz+1)
    integer :: i, j, k
    do k = 0, nz + 1
       do j = 0, ny + 1
          do i = 0, nx + 1
             PHIT(i, j, k) = phip(i, j, k) + PHI(i, j, k)
          end do
       end do
    end do
  end subroutine make_phit

  real(8) function phip(i, j, k)
    integer, intent(in) :: i, j, k
    integer :: s
    real(8) :: r
    phip = 0d0
    do s = 1, nmasses
       r = sqrt((masses(1, s) - xc(i))**2 + (masses(2, s) - yc(j))**2 + &
            (masses(3, s) - zc(k))**2 + softening**2)
       phip = phip - 6.67430d-8 * masses(4, s) / r
    end do
  end function phip

  ! calc_prims on the host (source/HydroCore.f90:14-21), for dry runs.
  subroutine host_prims(UPrim, U)
    real(8), intent(out) :: UPrim(5, -1:nx+2, -1:ny+2, -1:nz+2)
    real(8), intent(in) :: U(5, -1:nx+2, -1:ny+2, -1:nz+2)
    integer :: i, j, k
    do k = -1, nz + 2
       do j = -1, ny + 2
          do i = -1, nx + 2
             UPrim(1, i, j, k) = U(1, i, j, k)
             UPrim(2:4, i, j, k) = U(2:4, i, j, k) / U(1, i, j, k)
             UPrim(5, i, j, k) = max((gamma - 1d0) * (U(5, i, j, k) - 0.5d0 * &
                  (U(2, i, j, k)**2 + U(3, i, j, k)**2 + U(4, i, j, k)**2) / U(1, i, j, k)), 1d-20)
          end do
       end do
    end do
  end subroutine host_prims

  ! loaddata's format (source/cooling.f90:21-45): a row count, then rows of
  ! T and -log10(Lambda).
  subroutine load_table(path, temp, logl, n)
    character(*), intent(in) :: path
    real(8), allocatable, intent(out) :: temp(:), logl(:)
    integer, intent(out) :: n
    integer :: u, q
    open(newunit=u, file=trim(path), status='old', action='read')
    read(u, *) n
    allocate(temp(n), logl(n))
    do q = 1, n
       read(u, *) temp(q), logl(q)
       logl(q) = -logl(q)
    end do
    close(u)
  end subroutine load_table

end module driver_setup


module driver_output
  use driver_setup
  implicit none
contains

  ! Output number nout of UPrim, in the reference's formats
  ! (source/HydroCore.f90:253-395, single rank), and its sidecar
  ! <file>.time holding t, it and nout; then nout += 1 and tout = dtout * nout.
  subroutine output(UPrim, nout, tout, dtout, t, it, PHIT)
    real(8), intent(in) :: UPrim(5, -1:nx+2, -1:ny+2, -1:nz+2)
    integer, intent(inout) :: nout
    real(8), intent(inout) :: tout
    real(8), intent(in) :: dtout, t
    integer(c_int), intent(in) :: it
    real(8), intent(in), optional :: PHIT(0:nx+1, 0:ny+1, 0:nz+1)
    integer :: ios
    character(len=600) :: fname
    character(len=200) :: line
    character(len=1), parameter :: lf = char(10)
    integer :: u, i, j, k
    select case (trim(outfile))
    case ('bin')
       write(fname, '(a,a,i3.3,a,i3.3,a)') trim(outputpath), 'points', 0, 'N', nout, '.bin'
       open(newunit=u, file=trim(fname), status='replace', access='stream', &
            form='unformatted', convert='LITTLE_ENDIAN')
       write(u) UPrim
       close(u)
       if (present(PHIT)) then
          open(newunit=u, file='phi'//trim(fname), status='replace', access='stream', &
               form='unformatted', convert='LITTLE_ENDIAN', iostat=ios)
          if (ios == 0) then
             write(u) PHIT
             close(u)
          else
             write(*, '(a,a)') ' cannot open ', 'phi'//trim(fname)
          end if
       end if
    case ('dat')
       write(fname, '(a,a,i3.3,a,i3.3,a)') trim(outputpath), 'points', 0, 'N', nout, '.dat'
       open(newunit=u, file=trim(fname), status='replace')
       do k = 1, nz
          do j = 1, ny
             do i = 1, nx
                if (present(PHIT)) then
                   write(u, '(14(e16.7e3))') dble(i) * dx, dble(j) * dy, dble(k) * dz, &
                        UPrim(:, i, j, k), PHIT(i, j, k)
                else
                   write(u, '(14(e16.7e3))') dble(i) * dx, dble(j) * dy, dble(k) * dz, UPrim(:, i, j, k)
                end if
             end do
          end do
       end do
       close(u)
    case ('vtk')
       write(fname, '(a,a,i3.3,a,i3.3,a)') trim(outputpath), 'EPV', 0, 'N', nout, '.vtk'
       open(newunit=u, file=trim(fname), status='replace', access='stream', &
            form='unformatted', convert='BIG_ENDIAN')
       write(u) '# vtk DataFile Version 2.0', lf
       write(u) 'output from Diable', lf
       write(u) 'BINARY', lf
       write(u) 'DATASET STRUCTURED_POINTS', lf
       write(line, '("DIMENSIONS ",(3i6,1x))') nx, ny, nz
       write(u) trim(line), lf
       write(line, '("ORIGIN ",3e15.7)') xl, yl, zl
       write(u) trim(line), lf
       write(line, '("SPACING",3e15.7)') dx, dy, dz
       write(u) trim(line), lf
       write(line, '(a,i10)') 'POINT_DATA ', nx * ny * nz
       write(u) trim(line), lf
       write(u) 'SCALARS Density double 1', lf
       write(u) 'LOOKUP_TABLE default', lf
       write(u) (((UPrim(1, i, j, k), i = 1, nx), j = 1, ny), k = 1, nz)
       write(u) lf
       write(u) 'SCALARS Pressure double 1', lf
       write(u) 'LOOKUP_TABLE default', lf
       write(u) (((UPrim(5, i, j, k), i = 1, nx), j = 1, ny), k = 1, nz)
       if (present(PHIT)) then
          write(u) lf
          write(u) 'SCALARS Potential double 1', lf
          write(u) 'LOOKUP_TABLE default', lf
          write(u) (((PHIT(i, j, k), i = 1, nx), j = 1, ny), k = 1, nz)
       end if
       write(u) lf
       write(u) 'VECTORS Velocity double', lf
       write(u) (((UPrim(2, i, j, k), UPrim(3, i, j, k), UPrim(4, i, j, k), i = 1, nx), &
            j = 1, ny), k = 1, nz)
       write(u) lf
       close(u)
    case default
       write(*, '(a,a)') 'unknown outfile ', trim(outfile)
       error stop 2
    end select
    open(newunit=u, file=trim(fname)//'.time', status='replace')
    write(u, '(es25.17,1x,i0,1x,i0)') t, it, nout
    close(u)
    write(*, '(a,i0,a)') ' File ', nout, ' written'
    nout = nout + 1
    tout = dtout * nout
  end subroutine output

end module driver_output


program tlq_hydro
  use iso_c_binding
  use tlq_c
  use driver_setup
  use driver_output
  implicit none

  character(len=512) :: path
  real(8), allocatable :: U(:, :, :, :), UPrim(:, :, :, :), temp(:), logl(:)
  real(8), allocatable :: PHI(:, :, :), PHIT(:, :, :)
  type(tlq_params_t) :: prm
  type(c_ptr) :: ctx
  real(8) :: t, dt, tout, dtout
  integer(c_int) :: it, cit
  integer :: nout, ntab, u_phi, i, j, k
  logical :: phi_dump
  integer(c_long_long) :: bad

  if (command_argument_count() < 1) then
     write(*, '(a)') 'usage: tlq_hydro SETTINGS.nml'
     error stop 2
  end if
  call get_command_argument(1, path)
  call read_settings(trim(path))

  allocate(U(5, -1:nx+2, -1:ny+2, -1:nz+2), UPrim(5, -1:nx+2, -1:ny+2, -1:nz+2))
  U = 0d0
  call initflow(U)
  allocate(PHI(0:nx+1, 0:ny+1, 0:nz+1), PHIT(0:nx+1, 0:ny+1, 0:nz+1))
  PHI = 0d0
  if (len_trim(phifile) > 0) then
     open(newunit=u_phi, file=trim(phifile), status='old', access='stream', form='unformatted', &
          convert='LITTLE_ENDIAN', action='read')
     read(u_phi) PHI
     close(u_phi)
  end if
  if (trim(problem) == 'restart' .and. self_gravity /= TLQ_SG_OFF .and. .not. restart_fill) then
     inquire(file='phi'//trim(restart_file), exist=phi_dump)
     if (phi_dump) then
        open(newunit=u_phi, file='phi'//trim(restart_file), status='old', access='stream', &
             form='unformatted', convert='LITTLE_ENDIAN', action='read')
        read(u_phi) PHIT
        close(u_phi)
        do k = 0, nz + 1
           do j = 0, ny + 1
              do i = 0, nx + 1
                 PHI(i, j, k) = PHIT(i, j, k) - phip(i, j, k)
              end do
           end do
        end do
        write(*, '(a,a)') ' PHI from ', 'phi'//trim(restart_file)
     end if
  end if

  t = 0d0
  dt = 0d0
  it = 0
  nout = 0
  tout = 0d0
  dtout = tfin / max(num_out, 1)
  if (trim(problem) == 'restart') then
     call restart_clock(t, it, nout)
     nout = nout + 1
     tout = dtout * nout
     write(*, '(a,a,a,es14.6,a,i0)') ' Restart from ', trim(restart_file), ' t = ', t, ' it = ', it
  end if

  if (dry_run) then
     call host_prims(UPrim, U)
     if (self_gravity /= TLQ_SG_OFF) then
        call make_phit(PHI, PHIT)
        call output(UPrim, nout, tout, dtout, t, it, PHIT)
     else
        call output(UPrim, nout, tout, dtout, t, it)
     end if
  else
     call run_gpu()
  end if

contains

  subroutine run_gpu()
    prm = make_params()
    call tlq_require(tlq_create(prm, ctx), 'tlq_create')
    if (cooling /= TLQ_COOL_OFF) then
       call load_table(trim(cooling_file), temp, logl, ntab)
       call tlq_require(tlq_set_cooling_table(ctx, temp, logl, int(ntab, c_int)), &
            'tlq_set_cooling_table')
    end if
    if (nwinds > 0) call tlq_require(tlq_set_winds(ctx, winds, int(nwinds, c_int)), 'tlq_set_winds')
    if (nmasses > 0) call tlq_require(tlq_set_point_masses(ctx, masses, int(nmasses, c_int)), &
         'tlq_set_point_masses')
    call tlq_require(tlq_upload(ctx, TLQ_BUF_U, U), 'tlq_upload')
    if (self_gravity /= TLQ_SG_OFF) call tlq_require(tlq_upload_phi(ctx, PHI), 'tlq_upload_phi')
    if (per_routine) then
       if (trim(problem) /= 'restart' .or. restart_fill) &
            call tlq_require(tlq_borders(ctx, TLQ_BUF_U, 1_c_int), 'borders')     ! mainHD.f90:38
       call tlq_require(tlq_calc_prims(ctx, TLQ_BUF_U), 'calc_prims')             ! mainHD.f90:44
    else if (trim(problem) == 'restart') then
       call tlq_require(tlq_resume(ctx, t, it, merge(1_c_int, 0_c_int, restart_fill)), &
            'tlq_resume')
    else
       call tlq_require(tlq_init_state(ctx), 'tlq_init_state')
    end if

    do while (t < tfin .and. (max_steps <= 0 .or. it < max_steps))
       write(*, '(a,es14.6,a,es14.6,a,i0)') ' dt = ', dt, ' t = ', t, ' it = ', it
       if (t >= tout) call write_output()
       if (per_routine) then
          call tlq_require(tlq_time_step(ctx, dt), 'time_step')                   ! :74
          call tlq_require(tlq_fluxes(ctx, TLQ_BUF_U, 1_c_int), 'fluxes 1')         ! :76
          call tlq_require(tlq_godunov(ctx, TLQ_BUF_UPP, dt / 2d0), 'godunov UPP')  ! :78-80
          call tlq_require(tlq_borders(ctx, TLQ_BUF_UPP, 2_c_int), 'borders UPP')   ! :82
          call tlq_require(tlq_calc_prims(ctx, TLQ_BUF_UPP), 'calc_prims UPP')      ! :84
          call tlq_require(tlq_fluxes(ctx, TLQ_BUF_UPP, 2_c_int), 'fluxes 2')       ! :86
          call tlq_require(tlq_godunov(ctx, TLQ_BUF_UP, dt), 'godunov UP')          ! :88
          call tlq_require(tlq_borders(ctx, TLQ_BUF_UP, 1_c_int), 'borders UP')     ! :90
          call tlq_require(tlq_step(ctx, t, it, dt), 'step')                        ! :92
          call tlq_require(tlq_calc_prims(ctx, TLQ_BUF_U), 'calc_prims U')          ! :94
          if (cooling /= TLQ_COOL_OFF) call tlq_require(tlq_cooling(ctx, dt), 'cooling')  ! :96
       else
          call tlq_require(tlq_advance(ctx, dt, t, cit), 'tlq_advance')
          it = cit
       end if
    end do
    call write_output()
    call tlq_require(tlq_check_state(ctx, bad), 'tlq_check_state')
    write(*, '(a)') ' Done'
    write(*, '(a,es14.6,a,es14.6,a,i0,a,i0)') ' dt = ', dt, ' t = ', t, ' it = ', it, &
         ' bad cells = ', bad
    call tlq_require(tlq_destroy(ctx), 'tlq_destroy')
  end subroutine run_gpu

  subroutine write_output()
    if (.not. per_routine) call tlq_require(tlq_calc_prims(ctx, TLQ_BUF_U), 'calc_prims')
    call tlq_require(tlq_download(ctx, TLQ_BUF_UPRIM, UPrim), 'tlq_download')
    if (self_gravity /= TLQ_SG_OFF) then
       call tlq_require(tlq_download_phi(ctx, PHI), 'tlq_download_phi')
       call make_phit(PHI, PHIT)
       call output(UPrim, nout, tout, dtout, t, it, PHIT)
    else
       call output(UPrim, nout, tout, dtout, t, it)
    end if
  end subroutine write_output

end program tlq_hydro
