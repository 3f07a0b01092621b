! pf_case_io.f90 -- the executables' file contract, for the drop-in Fortran hosts.
!
! Inputs: config/controlDict.txt (seven namelist groups, src/omp_parallel/lib/global.f90:
! 47-62) and the porosity CSV (header "m,n[,l]", then one "i, j, k, value" row per cell,
! src/omp_parallel/lib/grid.f90:36-47, 166-177, 300-311).  Outputs, written with the same
! records (list-directed REAL(8) items grouped per WRITE exactly as the reference groups
! them, the same edit descriptors for the VTK files) so that a case directory holds the
! same bytes as after a run of the reference built with -fdefault-real-8:
!   etc/grid.dat                 output_grid_2d/3d              (output.f90:44-61, 588-609)
!   etc/solution_uvp.dat         output_solution_post_2d/3d     (output.f90:91-200, 648-791)
!   etc/surface_profile.dat      (same routines)
!   etc/divergent.dat            output_divergent_2d/3d         (output.f90:202-242, 912-966)
!   <output_folder>/output_NNNNN.vtk, output_paraview.vtk
!                                output_paraview(_temp)_2d/3d   (output.f90:307-537, 793-1088)
! and the standard-output log lines of read_settings, grid_conditions*, the time loop and
! get_now_time (global.f90:66-90, grid.f90:60-72 and 194-206, utils.f90:7-16).
module pf_case_io
  use iso_c_binding, only: c_double
  implicit none

  integer, parameter :: rk = c_double

  type pf_settings
    real(rk) :: xnue, xlambda, density, width, height, depth, time
    real(rk) :: inlet_velocity, outlet_pressure, aoa
    integer :: istep_max, istep_out
    real(rk) :: thickness, threshold, radius, center_x, center_y, center_z
    logical :: nonslip
    character(len=50) :: output_folder, csv_file
    integer :: iter_max
    real(rk) :: relux_factor
    ! optional group &gpu_control (not in the reference's controlDict; defaults = the
    ! reference's fixed-count FP64 solve)
    logical :: gpu_group = .false.
    character(len=16) :: precision = 'fp64'      ! 'fp64' or 'mixed'
    character(len=16) :: solver = 'sor'          ! 'sor' or 'multigrid' (tolerance mode)
    real(rk) :: residual_tol = 0.0_rk            ! > 0: tolerance mode, max|r/ap| <= residual_tol
    integer :: inner_iters = 0, max_cycles = 0
    real(rk) :: inner_tol = 0.0_rk
  end type pf_settings

contains

  subroutine print_time_stamp()
    character(len=20) :: stamp
    integer :: values(8)
    call date_and_time(values=values)
    write(stamp, '(I4.4,"-",I2.2,"-",I2.2," ",I2.2,":",I2.2,":",I2.2)') &
      values(1), values(2), values(3), values(5), values(6), values(7)
    write(*,*) '# --- TIME: ', trim(stamp)
  end subroutine print_time_stamp

  ! The seven namelist groups, read in order, then echoed; then the optional eighth group
  ! &gpu_control, read only when the file holds it (an existing controlDict parses as
  ! before and its run is unchanged).
  subroutine read_control_dict(s)
    type(pf_settings), intent(out) :: s
    real(rk) :: xnue, xlambda, density, width, height, depth, time
    real(rk) :: inlet_velocity, outlet_pressure, AoA
    integer :: istep_max, istep_out
    real(rk) :: thickness, threshold, radius, center_x, center_y, center_z
    logical :: nonslip
    character(len=50) :: output_folder, csv_file
    integer :: iter_max
    real(rk) :: relux_factor
    integer :: unit, ios
    character(len=16) :: precision, solver
    real(rk) :: residual_tol, inner_tol
    integer :: inner_iters, max_cycles
    namelist /gpu_control/ precision, solver, residual_tol, inner_iters, max_cycles, inner_tol
    namelist /physical/ xnue, xlambda, density, width, height, depth, time
    namelist /physical/ inlet_velocity, outlet_pressure, AoA
    namelist /file_control/ istep_out
    namelist /grid_control/ istep_max
    namelist /porosity_control/ thickness, threshold, radius, center_x, center_y, center_z
    namelist /calculation_method/ nonslip
    namelist /directory_control/ output_folder, csv_file
    namelist /solver_control/ iter_max, relux_factor

    open(newunit=unit, file="config/controlDict.txt", status="old", action="read")
    read(unit, nml=physical)
    read(unit, nml=file_control)
    read(unit, nml=grid_control)
    read(unit, nml=porosity_control)
    read(unit, nml=calculation_method)
    read(unit, nml=directory_control)
    read(unit, nml=solver_control)
    precision = 'fp64'
    solver = 'sor'
    residual_tol = 0.0_rk
    inner_iters = 0
    max_cycles = 0
    inner_tol = 0.0_rk
    read(unit, nml=gpu_control, iostat=ios)
    if (ios > 0) then
      write(*,*) 'controlDict: the &gpu_control group could not be read'
      error stop 1
    end if
    close(unit)
    s%gpu_group = (ios == 0)
    if (s%gpu_group) then
      if (precision /= 'fp64' .and. precision /= 'mixed') then
        write(*,*) 'controlDict: &gpu_control precision must be fp64 or mixed'
        error stop 1
      end if
      if (solver /= 'sor' .and. solver /= 'multigrid') then
        write(*,*) 'controlDict: &gpu_control solver must be sor or multigrid'
        error stop 1
      end if
      s%precision = precision
      s%solver = solver
      s%residual_tol = residual_tol
      s%inner_iters = inner_iters
      s%max_cycles = max_cycles
      s%inner_tol = inner_tol
    end if

    s%xnue = xnue; s%xlambda = xlambda; s%density = density
    s%width = width; s%height = height; s%depth = depth; s%time = time
    s%inlet_velocity = inlet_velocity; s%outlet_pressure = outlet_pressure; s%aoa = AoA
    s%istep_max = istep_max; s%istep_out = istep_out
    s%thickness = thickness; s%threshold = threshold; s%radius = radius
    s%center_x = center_x; s%center_y = center_y; s%center_z = center_z
    s%nonslip = nonslip
    s%output_folder = output_folder; s%csv_file = csv_file
    s%iter_max = iter_max; s%relux_factor = relux_factor

    write(*,*) '#'
    write(*,*) '# --- Physical conditions'
    write(*,*) '# xnue =', xnue
    write(*,*) '# xlambda =', xlambda
    write(*,*) '# density =', density
    write(*,*) '# width =', width
    write(*,*) '# height =', height
    write(*,*) '# depth =', depth
    write(*,*) '# time =', time
    write(*,*) '# inlet_velocity =', inlet_velocity
    write(*,*) '# outlet_pressure =', outlet_pressure
    write(*,*) '# Angle of inlet_velocity (AoA) =', AoA
    write(*,*) '#'
    write(*,*) '# --- Porosity information'
    write(*,*) '# thickness =', thickness
    write(*,*) '# threshold =', threshold
    write(*,*) '# radius =', radius
    write(*,*) '#'
    write(*,*) '# --- Directory information'
    write(*,*) '# output_folder =', output_folder
    write(*,*) '# input_porosity_file =', csv_file
    write(*,*) '#'
    write(*,*) '# --- Solver information'
    write(*,*) '# SOR max iteration steps =', iter_max
    write(*,*) '# SOR reluxation factor =', relux_factor
    if (s%gpu_group) then
      write(*,*) '#'
      write(*,*) '# --- GPU control'
      write(*,*) '# precision = ', trim(s%precision)
      write(*,*) '# solver = ', trim(s%solver)
      write(*,*) '# residual_tol =', s%residual_tol
      write(*,*) '# inner_iters =', s%inner_iters, ' max_cycles =', s%max_cycles, ' inner_tol =', s%inner_tol
    end if
  end subroutine read_control_dict

  ! Porosity CSV -> raw interior values, i fastest (the row's own i, j, k place each value).
  subroutine read_porosity_csv(csv_file, dim, m, n, l, raw)
    character(len=*), intent(in) :: csv_file
    integer, intent(in) :: dim
    integer, intent(out) :: m, n, l
    real(rk), allocatable, intent(out) :: raw(:)
    integer :: unit, x, y, z, c, total
    real(rk) :: val
    open(newunit=unit, file=csv_file, form='formatted', status='old', action='read')
    if (dim == 2) then
      read(unit, *) m, n
      l = 1
    else
      read(unit, *) m, n, l
    end if
    total = m * n * l
    allocate(raw(total))
    raw = 0.0_rk
    do c = 1, total
      read(unit, *) x, y, z, val
      if (dim == 2) z = 1
      raw(x + m * ((y - 1) + n * (z - 1))) = val
    end do
    close(unit)
  end subroutine read_porosity_csv

  ! Grid spacing, time step and the "Grid conditions" echo (grid.f90:50-72, 182-206).
  subroutine grid_setup(s, dim, m, n, l, dx, dy, dz, dt, xp, yp, zp)
    type(pf_settings), intent(in) :: s
    integer, intent(in) :: dim, m, n, l
    real(rk), intent(out) :: dx, dy, dz, dt
    real(rk), allocatable, intent(out) :: xp(:), yp(:), zp(:)
    real(rk) :: cfl_no, pecret_no, diffusion_factor, reynolds_no
    integer :: i, j, k

    dx = s%width / real(m - 1, rk)
    dy = s%height / real(n - 1, rk)
    if (dim == 3) then
      dz = s%depth / real(l - 1, rk)
    else
      dz = 1.0_rk
    end if
    dt = s%time / real(s%istep_max, rk)
    cfl_no = s%inlet_velocity * dt / dx
    pecret_no = s%inlet_velocity * dx / s%xnue
    diffusion_factor = s%xnue * dt / dy / dy
    reynolds_no = s%inlet_velocity * s%radius * 2.0_rk / s%xnue

    write(*,*)
    write(*,*) '# --- Grid conditions'
    if (dim == 2) then
      write(*,*) '# m, n =', m, n
      write(*,*) '# dx, dy =', dx, dy
    else
      write(*,*) '# m, n, l =', m, n, l
      write(*,*) '# istep_max =', s%istep_max
      write(*,*) '# dx, dy, dz =', dx, dy, dz
    end if
    write(*,*) '# dt =', dt
    write(*,*) '# cfl_no =', cfl_no
    write(*,*) '# pecret_no =', pecret_no
    write(*,*) '# diffusion_factor =', diffusion_factor
    write(*,*) '# reynolds_no =', reynolds_no
    if (dim == 3) then
      write(*,*) '# thickness =', s%thickness
      write(*,*) '# threshold =', s%threshold
    end if
    write(*,*)

    ! cell-centre coordinates; zp spans at least 0:m+1 because the 3D surface profile
    ! prints zp(i) (output.f90:781), which the reference reads from its zero-initialised
    ! static array beyond l+1
    allocate(xp(0:m + 1), yp(0:n + 1), zp(0:max(m, l) + 1))
    zp = 0.0_rk
    do i = 0, m + 1
      xp(i) = dx * real(i - 1, rk) - s%width * s%center_x
    end do
    do j = 0, n + 1
      yp(j) = dy * real(j - 1, rk) - s%height * s%center_y
    end do
    if (dim == 3) then
      do k = 0, l + 1
        zp(k) = dz * real(k - 1, rk) - s%depth * s%center_z
      end do
    end if
  end subroutine grid_setup

  subroutine write_grid(dim, m, n, l, xp, yp, zp)
    integer, intent(in) :: dim, m, n, l
    real(rk), intent(in) :: xp(0:), yp(0:), zp(0:)
    integer :: unit, i, j, k
    open(newunit=unit, file='etc/grid.dat', status='replace')
    if (dim == 2) then
      write(unit,*) 'm, n =', m, n
    else
      write(unit,*) 'm, n, l =', m, n, l
    end if
    write(unit,*) 'grid points ='
    write(unit,*) (xp(i), i=1,m)
    write(unit,*) (yp(j), j=1,n)
    if (dim == 3) write(unit,*) (zp(k), k=1,l)
    close(unit)
  end subroutine write_grid

  ! ---------------------------------------------------------------------------------------
  ! 2D writers; fields are (0:m+1, 0:n+1)
  ! ---------------------------------------------------------------------------------------
  subroutine write_solution_post_2d(m, n, p, u, v, po, xp, yp)
    integer, intent(in) :: m, n
    real(rk), intent(in) :: p(0:,0:), u(0:,0:), v(0:,0:), po(0:,0:), xp(0:), yp(0:)
    real(rk), parameter :: small = 1.0e-6_rk, zero = 0.0_rk, pmin = 0.25_rk, pmax = 0.75_rk
    real(rk), allocatable :: uc(:,:), vc(:,:), pc(:,:)
    integer :: unit, i, j
    allocate(uc(m, n), vc(m, n), pc(m, n))
    do j = 1, n
      do i = 1, m
        uc(i,j) = u(i,j) * po(i,j)
        vc(i,j) = v(i,j) * po(i,j)
        if (po(i,j) > small) then
          pc(i,j) = p(i,j)
        else
          pc(i,j) = zero
        end if
      end do
    end do
    open(newunit=unit, file='etc/solution_uvp.dat', status='replace')
    write(unit,*) 'm, n =', m, n
    write(unit,*) 'velocity u_bulk '
    do j = 1, n
      write(unit,*) (uc(i,j), i=1,m)
    end do
    write(unit,*) 'velocity v_bulk '
    do j = 1, n
      write(unit,*) (vc(i,j), i=1,m)
    end do
    write(unit,*) 'velocity u_inst '
    do j = 1, n
      write(unit,*) (u(i,j), i=1,m)
    end do
    write(unit,*) 'velocity v_inst '
    do j = 1, n
      write(unit,*) (v(i,j), i=1,m)
    end do
    write(unit,*) 'pressure p_fluid'
    do j = 1, n
      write(unit,*) (pc(i,j), i=1,m)
    end do
    write(unit,*) 'pressure P_all'
    do j = 1, n
      write(unit,*) (p(i,j), i=1,m)
    end do
    write(unit,*) 'porosity'
    do j = 1, n
      write(unit,*) (po(i,j), i=1,m)
    end do
    close(unit)

    open(newunit=unit, file='etc/surface_profile.dat', status='replace')
    do j = 1, n
      do i = 1, m
        if (po(i,j) < pmax .and. po(i,j) > pmin) write(unit,*) xp(i), yp(j), pc(i,j), po(i,j)
      end do
    end do
    close(unit)
  end subroutine write_solution_post_2d

  ! div: interior values, i fastest (pf_divergence)
  subroutine write_divergent_2d(m, n, po, div)
    integer, intent(in) :: m, n
    real(rk), intent(in) :: po(0:,0:), div(:)
    integer :: unit, i, j
    open(newunit=unit, file='etc/divergent.dat', status='replace')
    write(unit,*)
    write(unit,*) 'porosity'
    do j = 1, n
      write(unit,*) (po(i,j), i=1,m)
    end do
    write(unit,*)
    write(unit,*) 'divergent velocity'
    do j = 1, n
      write(unit,*) (div(i + m * (j - 1)), i=1,m)
    end do
    write(unit,*)
    close(unit)
  end subroutine write_divergent_2d

  ! output_paraview_2d / output_paraview_temp_2d: both write the same arrays.
  subroutine write_vtk_2d(path, m, n, p, u, v, po, xp, yp, uin)
    character(len=*), intent(in) :: path
    integer, intent(in) :: m, n
    real(rk), intent(in) :: p(0:,0:), u(0:,0:), v(0:,0:), po(0:,0:), xp(0:), yp(0:), uin
    real(rk) :: dv
    integer :: unit, i, j
    open(newunit=unit, file=path, status="unknown", form="formatted", position="rewind")
    write(unit,"('# vtk DataFile Version 3.0')")
    write(unit,"('2D flow')")
    write(unit,"('ASCII ')")
    write(unit,"('DATASET STRUCTURED_GRID')")
    write(unit,"('DIMENSIONS ',3(1x,i4))") m, n, 1
    write(unit,"('POINTS ',i9,' float')") m*n
    do j = 1, n
      do i = 1, m
        write(unit,"(3(f16.4,1x))") xp(i), yp(j), 0.0_rk
      end do
    end do
    write(unit,"('POINT_DATA ',i9)") m*n
    write(unit,"('VECTORS velocity float')")
    do j = 1, n
      do i = 1, m
        write(unit,"(3(f16.4,1x))") u(i,j), v(i,j), 0.0_rk
      end do
    end do
    write(unit,"('VECTORS velocityInFluid float')")
    do j = 1, n
      do i = 1, m
        write(unit,"(3(f16.4,1x))") u(i,j)*po(i,j), v(i,j)*po(i,j), 0.0_rk
      end do
    end do
    write(unit,"('VECTORS dimless_v float')")
    do j = 1, n
      do i = 1, m
        write(unit,"(3(f16.4,1x))") u(i,j)*po(i,j)/uin, v(i,j)*po(i,j)/uin, 0.0_rk
      end do
    end do
    write(unit,"('SCALARS porosity float')")
    write(unit,"('LOOKUP_TABLE default')")
    do j = 1, n
      do i = 1, m
        write(unit,"(3(f16.4,1x))") po(i,j)
      end do
    end do
    write(unit,"('SCALARS pressure float')")
    write(unit,"('LOOKUP_TABLE default')")
    do j = 1, n
      do i = 1, m
        write(unit,"(3(f16.4,1x))") p(i,j)
      end do
    end do
    write(unit,"('SCALARS VelocityDivergent float')")
    write(unit,"('LOOKUP_TABLE default')")
    do j = 1, n
      do i = 1, m
        dv = (u(i+1,j)-u(i-1,j))/(xp(i+1)-xp(i-1)) + (v(i,j+1)-v(i,j-1))/(yp(j+1)-yp(j-1))
        write(unit,"(3(f16.4,1x))") dv
      end do
    end do
    write(unit,"('SCALARS abs_dimless_v float')")
    write(unit,"('LOOKUP_TABLE default')")
    do j = 1, n
      do i = 1, m
        write(unit,"(3(f16.4,1x))") sqrt((u(i,j)*po(i,j)/uin)**2 + (v(i,j)*po(i,j)/uin)**2)
      end do
    end do
    close(unit)
  end subroutine write_vtk_2d

  ! ---------------------------------------------------------------------------------------
  ! 3D writers; fields are (0:m+1, 0:n+1, 0:l+1)
  ! ---------------------------------------------------------------------------------------
  subroutine write_solution_post_3d(m, n, l, p, u, v, w, po, xp, yp, zp)
    integer, intent(in) :: m, n, l
    real(rk), intent(in) :: p(0:,0:,0:), u(0:,0:,0:), v(0:,0:,0:), w(0:,0:,0:), po(0:,0:,0:)
    real(rk), intent(in) :: xp(0:), yp(0:), zp(0:)
    real(rk), parameter :: small = 1.0e-6_rk, zero = 0.0_rk, pmin = 0.25_rk, pmax = 0.75_rk
    real(rk), allocatable :: uc(:,:,:), vc(:,:,:), wc(:,:,:), pc(:,:,:)
    integer :: unit, i, j, k
    allocate(uc(m, n, l), vc(m, n, l), wc(m, n, l), pc(m, n, l))
    do k = 1, l
      do j = 1, n
        do i = 1, m
          uc(i,j,k) = u(i,j,k) * po(i,j,k)
          vc(i,j,k) = v(i,j,k) * po(i,j,k)
          wc(i,j,k) = w(i,j,k) * po(i,j,k)
          if (po(i,j,k) > small) then
            pc(i,j,k) = p(i,j,k)
          else
            pc(i,j,k) = zero
          end if
        end do
      end do
    end do
    open(newunit=unit, file='etc/solution_uvp.dat', status='replace')
    write(unit,*) 'm, n, l =', m, n, l
    write(unit,*) 'velocity u_bulk '
    do k = 1, l
      write(unit,*) ((uc(i,j,k), i=1,m), j=1,n)
    end do
    write(unit,*) 'velocity v_bulk '
    do k = 1, l
      write(unit,*) ((vc(i,j,k), i=1,m), j=1,n)
    end do
    write(unit,*) 'velocity w_bulk '
    do k = 1, l
      write(unit,*) ((wc(i,j,k), i=1,m), j=1,n)
    end do
    write(unit,*) 'velocity u_inst '
    do k = 1, l
      write(unit,*) ((u(i,j,k), i=1,m), j=1,n)
    end do
    write(unit,*) 'velocity v_inst '
    do k = 1, l
      write(unit,*) ((v(i,j,k), i=1,m), j=1,n)
    end do
    write(unit,*) 'velocity w_inst '
    do k = 1, l
      write(unit,*) ((w(i,j,k), i=1,m), j=1,n)
    end do
    write(unit,*) 'pressure p_fluid'
    do k = 1, l
      write(unit,*) ((pc(i,j,k), i=1,m), j=1,n)
    end do
    write(unit,*) 'pressure P_all'
    do k = 1, l
      write(unit,*) ((p(i,j,k), i=1,m), j=1,n)
    end do
    write(unit,*) 'porosity'
    do k = 1, l
      write(unit,*) ((po(i,j,k), i=1,m), j=1,n)
    end do
    close(unit)

    open(newunit=unit, file='etc/surface_profile.dat', status='replace')
    do k = 1, l
      do j = 1, n
        do i = 1, m
          ! zp(i), not zp(k): the reference's record (output.f90:781)
          if (po(i,j,k) < pmax .and. po(i,j,k) > pmin) write(unit,*) xp(i), yp(j), zp(i), pc(i,j,k), po(i,j,k)
        end do
      end do
    end do
    close(unit)
  end subroutine write_solution_post_3d

  subroutine write_divergent_3d(m, n, l, po, div)
    integer, intent(in) :: m, n, l
    real(rk), intent(in) :: po(0:,0:,0:), div(:)
    integer :: unit, i, j, k
    open(newunit=unit, file='etc/divergent.dat', status='replace')
    write(unit,*)
    write(unit,*) 'porosity'
    do k = 1, l
      do j = 1, n
        write(unit,*) (po(i,j,k), i=1,m)
      end do
    end do
    write(unit,*)
    write(unit,*) 'divergent velocity'
    do k = 1, l
      do j = 1, n
        write(unit,*) (div(i + m * ((j - 1) + n * (k - 1))), i=1,m)
      end do
    end do
    write(unit,*)
    close(unit)
  end subroutine write_divergent_3d

  ! final = .true.: output_paraview_3d (velocity, velocityInFluid, pressure,
  ! VelocityDivergent, porosity); .false.: output_paraview_temp_3d (velocity,
  ! velocityInFluid, porosity, pressure, VelocityDivergent).
  subroutine write_vtk_3d(path, final, m, n, l, p, u, v, w, po, xp, yp, zp)
    character(len=*), intent(in) :: path
    logical, intent(in) :: final
    integer, intent(in) :: m, n, l
    real(rk), intent(in) :: p(0:,0:,0:), u(0:,0:,0:), v(0:,0:,0:), w(0:,0:,0:), po(0:,0:,0:)
    real(rk), intent(in) :: xp(0:), yp(0:), zp(0:)
    integer :: unit, i, j, k
    open(newunit=unit, file=path, status="unknown", form="formatted", position="rewind")
    write(unit,"('# vtk DataFile Version 3.0')")
    write(unit,"('3D flow')")
    write(unit,"('ASCII ')")
    write(unit,"('DATASET STRUCTURED_GRID')")
    write(unit,"('DIMENSIONS ',3(1x,i4))") m, n, l
    write(unit,"('POINTS ',i9,' float')") m*n*l
    do k = 1, l
      do j = 1, n
        do i = 1, m
          write(unit,"(3(f16.4,1x))") xp(i), yp(j), zp(k)
        end do
      end do
    end do
    write(unit,"('POINT_DATA ',i9)") m*n*l
    write(unit,"('VECTORS velocity float')")
    do k = 1, l
      do j = 1, n
        do i = 1, m
          write(unit,"(3(f16.4,1x))") u(i,j,k), v(i,j,k), w(i,j,k)
        end do
      end do
    end do
    write(unit,"('VECTORS velocityInFluid float')")
    do k = 1, l
      do j = 1, n
        do i = 1, m
          write(unit,"(3(f16.4,1x))") u(i,j,k)*po(i,j,k), v(i,j,k)*po(i,j,k), w(i,j,k)*po(i,j,k)
        end do
      end do
    end do
    if (final) then
      call scalars_p()
      call scalars_div()
      call scalars_po()
    else
      call scalars_po()
      call scalars_p()
      call scalars_div()
    end if
    close(unit)
  contains
    subroutine scalars_p()
      write(unit,"('SCALARS pressure float')")
      write(unit,"('LOOKUP_TABLE default')")
      do k = 1, l
        do j = 1, n
          do i = 1, m
            write(unit,"(3(f16.4,1x))") p(i,j,k)
          end do
        end do
      end do
    end subroutine scalars_p
    subroutine scalars_po()
      write(unit,"('SCALARS porosity float')")
      write(unit,"('LOOKUP_TABLE default')")
      do k = 1, l
        do j = 1, n
          do i = 1, m
            write(unit,"(3(f16.4,1x))") po(i,j,k)
          end do
        end do
      end do
    end subroutine scalars_po
    subroutine scalars_div()
      real(rk) :: dv
      write(unit,"('SCALARS VelocityDivergent float')")
      write(unit,"('LOOKUP_TABLE default')")
      do k = 1, l
        do j = 1, n
          do i = 1, m
            dv = (u(i+1,j,k)-u(i-1,j,k))/(xp(i+1)-xp(i-1)) &
               + (v(i,j+1,k)-v(i,j-1,k))/(yp(j+1)-yp(j-1)) &
               + (w(i,j,k+1)-w(i,j,k-1))/(zp(k+1)-zp(k-1))
            write(unit,"(3(f16.4,1x))") dv
          end do
        end do
      end do
    end subroutine scalars_div
  end subroutine write_vtk_3d

end module pf_case_io
