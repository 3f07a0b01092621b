! pf_driver.f90 -- the time loop of the drop-in executables over the C-ABI.
!
! One routine serves the five reference executables; the variant selects what differs:
!   '2d_uniform'   ibm_2d_uniform_omp_cpu.f90 program main (:28-136)
!   '2d_drag'      ibm_2d_drag_omp_cpu.f90 (the same step, force log every step, :121)
!   '2d_backstep'  ibm_2d_backstep_omp_cpu.f90 (inlet/initial velocity times porosity)
!   '3d_uniform'   ibm_3d_uniform_omp_cpu.f90 program main (:31-144)
!   '3d_air'       ibm_3d_air_condition_omp_cpu.f90 program main (:45-158), face codes of
!                  its wall_conditions module (:10-15)
! The fields live on the GPU for the whole run; the host downloads them only for the VTK
! snapshots every istep_out steps and the final outputs.  Steps between two snapshots are
! queued in one pf_step call and their log lines printed from its statistics afterwards,
! in the reference's order (the drag variant steps one at a time, its force log follows
! each step).
module pf_driver
  use iso_c_binding
  use pf_binding
  use pf_case_io
  implicit none

contains

  subroutine run_case(variant)
    character(len=*), intent(in) :: variant
    type(pf_settings) :: s
    type(pf_desc) :: d
    type(c_ptr) :: ctx
    integer :: dim, m, n, l, istep, chunk, q, st
    real(rk) :: dx, dy, dz, dt, time
    real(rk), allocatable :: raw(:), xp(:), yp(:), zp(:), div(:)
    real(rk), allocatable :: u(:,:,:), v(:,:,:), w(:,:,:), p(:,:,:), po(:,:,:)
    type(pf_stats), allocatable :: stats(:)
    real(rk) :: fr(6)
    logical :: drag, air
    integer(c_int64_t) :: ldi, ldj

    drag = (variant == '2d_drag')
    air = (variant == '3d_air')
    if (variant(1:2) == '2d') then
      dim = 2
    else
      dim = 3
    end if
    if (pf_abi_version() /= PF_ABI_VERSION_EXPECTED) then
      write(*,*) 'libpixelflow ABI version', pf_abi_version(), ' expected', PF_ABI_VERSION_EXPECTED
      error stop 1
    end if

    call print_time_stamp()
    call read_control_dict(s)
    call execute_command_line('mkdir -p '//trim(s%output_folder))
    call execute_command_line('mkdir -p etc')
    call read_porosity_csv(trim(s%csv_file), dim, m, n, l, raw)
    call grid_setup(s, dim, m, n, l, dx, dy, dz, dt, xp, yp, zp)
    call write_grid(dim, m, n, l, xp, yp, zp)
    istep = 0
    write(*,*) '# istep_max= ', s%istep_max, '   istep_out= ', s%istep_out

    d%dim = dim
    d%m = m
    d%n = n
    d%l = l
    if (air) then
      d%bc = PF_BC_AIR
      d%face = [1, 0, 0, 0, 0, 2]   ! top inlet, south outlet, the rest walls
    else
      d%bc = PF_BC_UNIFORM
    end if
    if (s%nonslip) d%nonslip = 1
    d%iter_max = s%iter_max
    d%precision = PF_PREC_FP64
    if (s%precision == 'mixed') d%precision = PF_PREC_MIXED
    if (s%solver == 'multigrid') d%solver = PF_SOLVER_MG
    d%residual_tol = s%residual_tol
    d%inner_iters = s%inner_iters
    d%max_cycles = s%max_cycles
    d%inner_tol = s%inner_tol
    if (variant == '2d_backstep') d%inlet_porosity = 1
    d%dx = dx
    d%dy = dy
    d%dz = dz
    d%dt = dt
    d%xnue = s%xnue
    d%xlambda = s%xlambda
    d%density = s%density
    d%thickness = s%thickness
    d%inlet_velocity = s%inlet_velocity
    d%outlet_pressure = s%outlet_pressure
    d%aoa = s%aoa
    d%omega = s%relux_factor

    ctx = c_null_ptr
    call pf_check(pf_create(d, ctx), ctx, 'pf_create')
    call pf_check(pf_set_porosity(ctx, raw, s%threshold), ctx, 'pf_set_porosity')
    call pf_check(pf_initial(ctx), ctx, 'pf_initial')

    allocate(u(0:m+1, 0:n+1, 0:l+1), v(0:m+1, 0:n+1, 0:l+1), p(0:m+1, 0:n+1, 0:l+1), &
             po(0:m+1, 0:n+1, 0:l+1), w(0:m+1, 0:n+1, 0:l+1))
    if (dim == 2) then
      deallocate(u, v, p, po, w)
      allocate(u(0:m+1, 0:n+1, 0:0), v(0:m+1, 0:n+1, 0:0), p(0:m+1, 0:n+1, 0:0), &
               po(0:m+1, 0:n+1, 0:0), w(0:m+1, 0:n+1, 0:0))
    end if
    u = 0.0_rk
    v = 0.0_rk
    w = 0.0_rk
    p = 0.0_rk
    ldi = int(m + 2, c_int64_t)
    ldj = int(n + 2, c_int64_t)
    call pf_check(pf_download(ctx, PF_FIELD_POROSITY, po, ldi, ldj), ctx, 'pf_download porosity')
    call snapshot(0)

    call print_time_stamp()
    write(*,*) '# --- MAC algorithm start'

    allocate(stats(max(1, s%istep_out)))
    istep = 0
    do while (istep < s%istep_max)
      ! steps up to the next snapshot (or the end) in one call
      if (drag) then
        chunk = 1
      else
        chunk = min(s%istep_out - mod(istep, s%istep_out), s%istep_max - istep)
      end if
      if (chunk > size(stats)) then
        deallocate(stats)
        allocate(stats(chunk))
      end if
      st = pf_step(ctx, chunk, stats)
      if (st /= PF_OK .and. st /= PF_EDIVERGED) call pf_check(st, ctx, 'pf_step')
      do q = 1, chunk
        istep = istep + 1
        time = istep * dt
        write(*,*) '--- time_steps= ', istep, ' --  time = ', time
        if (s%residual_tol > 0.0_rk) then
          ! tolerance mode: the iterations the step ran (V-cycles for the multigrid solve)
          ! and the FP64 residual it reached
          write(*,*) 'SOR iteration no.', stats(q)%sor_iters, '-- p error:', stats(q)%sor_error, &
                     '-- residual:', stats(q)%residual
        else
          write(*,*) 'SOR iteration no.', s%iter_max, '-- p error:', stats(q)%sor_error
        end if
      end do
      if (drag) call force_log()
      if (mod(istep, s%istep_out) == 0) call snapshot(istep)
    end do

    call print_time_stamp()
    call fetch_fields()
    allocate(div(m * n * l))
    call pf_check(pf_divergence(ctx, div), ctx, 'pf_divergence')
    if (dim == 2) then
      call write_solution_post_2d(m, n, p(:,:,0), u(:,:,0), v(:,:,0), po(:,:,0), xp, yp)
      call write_divergent_2d(m, n, po(:,:,0), div)
      call write_vtk_2d(trim(s%output_folder)//'/output_paraview.vtk', m, n, p(:,:,0), u(:,:,0), v(:,:,0), &
                        po(:,:,0), xp, yp, s%inlet_velocity)
    else
      call write_solution_post_3d(m, n, l, p, u, v, w, po, xp, yp, zp)
      call write_divergent_3d(m, n, l, po, div)
      call write_vtk_3d(trim(s%output_folder)//'/output_paraview.vtk', .true., m, n, l, p, u, v, w, po, xp, yp, zp)
    end if
    write(*,*) 'program finished'
    call print_time_stamp()
    call pf_check(pf_destroy(ctx), c_null_ptr, 'pf_destroy')

  contains

    subroutine fetch_fields()
      call pf_check(pf_download(ctx, PF_FIELD_U, u, ldi, ldj), ctx, 'pf_download u')
      call pf_check(pf_download(ctx, PF_FIELD_V, v, ldi, ldj), ctx, 'pf_download v')
      if (dim == 3) call pf_check(pf_download(ctx, PF_FIELD_W, w, ldi, ldj), ctx, 'pf_download w')
      call pf_check(pf_download(ctx, PF_FIELD_P, p, ldi, ldj), ctx, 'pf_download p')
    end subroutine fetch_fields

    subroutine snapshot(k)
      integer, intent(in) :: k
      character(len=5) :: number
      write(number, "(I5.5)") k
      call fetch_fields()
      if (dim == 2) then
        call write_vtk_2d(trim(s%output_folder)//'/output_'//number//'.vtk', m, n, p(:,:,0), u(:,:,0), &
                          v(:,:,0), po(:,:,0), xp, yp, s%inlet_velocity)
      else
        call write_vtk_3d(trim(s%output_folder)//'/output_'//number//'.vtk', .false., m, n, l, p, u, v, w, po, &
                          xp, yp, zp)
      end if
    end subroutine snapshot

    ! output_force_log_2d's lines (output.f90:288-303) from the device reduction.
    subroutine force_log()
      real(rk) :: fx, fy, cd, cl
      call pf_check(pf_force(ctx, fr), ctx, 'pf_force')
      fx = fr(1) + fr(4)
      fy = fr(2) + fr(5)
      cd = fx / (s%density * s%inlet_velocity ** 2 * s%radius)
      cl = fy / (s%density * s%inlet_velocity ** 2 * s%radius)
      write(*,*) 'Fp =', fr(1), fr(2)
      write(*,*) 'Fv =', fr(4), fr(5)
      write(*,*) 'F  =', fx, fy
      write(*,*) 'Cd =', cd, 'Cl =', cl
    end subroutine force_log

  end subroutine run_case

end module pf_driver
