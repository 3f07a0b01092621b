! pf_binding.f90 -- ISO_C_BINDING interface of include/pixelflow.h for Fortran hosts.
!
! The derived types mirror pf_desc and pf_stats field by field (bind(C) gives them the C
! layout); the interfaces bind the single-case entry points.  Host arrays are passed as
! the reference keeps them: Fortran column-major with ghost layers, leading dimensions
! ldi = size(a, 1), ldj = size(a, 2).  Every call returns 0 or a negative PF_E* code;
! pf_check turns a failure into an error stop with the library's message.
module pf_binding
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: PF_OK = 0, PF_EINVAL = -1, PF_EHIP = -2, PF_ENOMEM = -3, &
                               PF_ESTATE = -4, PF_EDIVERGED = -5
  integer(c_int), parameter :: PF_BC_UNIFORM = 0, PF_BC_AIR = 1
  integer(c_int), parameter :: PF_PREC_FP64 = 0, PF_PREC_MIXED = 1
  integer(c_int), parameter :: PF_FIELD_U = 0, PF_FIELD_V = 1, PF_FIELD_W = 2, PF_FIELD_P = 3, &
                               PF_FIELD_POROSITY = 4
  integer(c_int), parameter :: PF_SOLVER_SOR = 0, PF_SOLVER_MG = 1
  integer(c_int), parameter :: PF_ABI_VERSION_EXPECTED = 3

  type, bind(C) :: pf_desc
    integer(c_int32_t) :: dim = 0, m = 0, n = 0, l = 1
    integer(c_int32_t) :: bc = 0, nonslip = 0, iter_max = 0
    integer(c_int32_t) :: face(6) = 0
    integer(c_int32_t) :: precision = 0, inner_iters = 0, max_cycles = 0, report_residual = 0
    integer(c_int32_t) :: inlet_porosity = 0
    integer(c_int32_t) :: solver = 0
    real(c_double) :: dx = 0, dy = 0, dz = 0, dt = 0
    real(c_double) :: xnue = 0, xlambda = 0, density = 0, thickness = 0
    real(c_double) :: inlet_velocity = 0, outlet_pressure = 0, aoa = 0
    real(c_double) :: omega = 0
    real(c_double) :: residual_tol = 0
    real(c_double) :: inner_tol = 0
  end type pf_desc

  type, bind(C) :: pf_stats
    real(c_double) :: sor_error = 0, residual = 0
    integer(c_int32_t) :: sor_iters = 0, cycles = 0
  end type pf_stats

  interface
    integer(c_int) function pf_abi_version() bind(C, name="pf_abi_version")
      import :: c_int
    end function pf_abi_version

    integer(c_int) function pf_create(desc, ctx) bind(C, name="pf_create")
      import :: c_int, c_ptr, pf_desc
      type(pf_desc), intent(in) :: desc
      type(c_ptr), intent(out) :: ctx
    end function pf_create

    integer(c_int) function pf_destroy(ctx) bind(C, name="pf_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function pf_destroy

    type(c_ptr) function pf_last_error(ctx) bind(C, name="pf_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function pf_last_error

    integer(c_int) function pf_set_porosity(ctx, raw, threshold) bind(C, name="pf_set_porosity")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: raw(*)
      real(c_double), value :: threshold
    end function pf_set_porosity

    integer(c_int) function pf_initial(ctx) bind(C, name="pf_initial")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function pf_initial

    integer(c_int) function pf_upload(ctx, field, host, ldi, ldj) bind(C, name="pf_upload")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: field
      real(c_double), intent(in) :: host(*)
      integer(c_int64_t), value :: ldi, ldj
    end function pf_upload

    integer(c_int) function pf_download(ctx, field, host, ldi, ldj) bind(C, name="pf_download")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: field
      real(c_double), intent(inout) :: host(*)
      integer(c_int64_t), value :: ldi, ldj
    end function pf_download

    integer(c_int) function pf_step(ctx, nsteps, stats) bind(C, name="pf_step")
      import :: c_int, c_ptr, pf_stats
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps
      type(pf_stats), intent(inout) :: stats(*)
    end function pf_step

    integer(c_int) function pf_divergence(ctx, out) bind(C, name="pf_divergence")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: out(*)
    end function pf_divergence

    integer(c_int) function pf_force(ctx, out) bind(C, name="pf_force")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: out(6)
    end function pf_force

    integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function c_strlen
  end interface

contains

  ! The library's message for the last failure of ctx (or of pf_create when ctx is null).
  function pf_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: i, nchar
    p = pf_last_error(ctx)
    if (.not. c_associated(p)) then
      msg = ''
      return
    end if
    nchar = int(c_strlen(p))
    call c_f_pointer(p, chars, [nchar])
    allocate(character(len=nchar) :: msg)
    do i = 1, nchar
      msg(i:i) = chars(i)
    end do
  end function pf_message

  subroutine pf_check(rc, ctx, what)
    integer(c_int), intent(in) :: rc
    type(c_ptr), intent(in) :: ctx
    character(len=*), intent(in) :: what
    if (rc /= PF_OK) then
      write(*, '(a, a, i0, a, a)') trim(what), ' failed (', rc, '): ', pf_message(ctx)
      error stop 1
    end if
  end subroutine pf_check

end module pf_binding
