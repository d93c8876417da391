! ISO_C_BINDING interface to libvanadis_gpu.so for a Fortran host.
!
! The legacy entry kernel_dbcg_simple_ needs no interface: the host's
! implicit-external CALL kernel_dbcg_simple(...) (VANADIS.FOR:809-810) already
! binds to it with every argument by reference.  This module declares the
! context API of include/vanadis_gpu.h for hosts that keep the operator and
! the solver state on the device across time steps.
module vanadis_gpu
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: VGPU_OK = 0
  integer(c_int), parameter :: VGPU_NOT_CONVERGED = 5
  integer(c_int), parameter :: VGPU_SCATTER_OWNER = 0, VGPU_SCATTER_COLORED = 1, &
                               VGPU_SCATTER_ATOMIC = 2
  integer(c_int), parameter :: VGPU_DR_V1 = 1, VGPU_DR_V2 = 2, VGPU_DR_V3 = 3
  integer(c_int), parameter :: VGPU_SCHEME_THETA = 0, VGPU_SCHEME_STEADY = 1, &
                               VGPU_SCHEME_EULER = 2
  integer(c_int), parameter :: VGPU_FP64 = 0, VGPU_FP32 = 1, VGPU_MIXED = 2
  integer(c_int), parameter :: VGPU_METHOD_BICG = 0, VGPU_METHOD_BICGSTAB = 1

  interface
    integer(c_int) function vgpu_create(nax, nay, naz, xx, yy, zz, iss, bc, ctx) &
        bind(C, name="vgpu_create")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: nax, nay, naz
      real(c_double), intent(in) :: xx(*), yy(*), zz(*)
      integer(c_int), intent(in) :: iss(*), bc(*)
      type(c_ptr), intent(out) :: ctx
    end function

    integer(c_int) function vgpu_destroy(ctx) bind(C, name="vgpu_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function vgpu_set_coeffs(ctx, lam, vel, p, alpha, variant, scheme, dt) &
        bind(C, name="vgpu_set_coeffs")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: lam(*), vel(*)
      real(c_double), value :: p, alpha, dt
      integer(c_int), value :: variant, scheme
    end function

    integer(c_int) function vgpu_set_options(ctx, scatter, zchunk, tile) &
        bind(C, name="vgpu_set_options")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: scatter, zchunk, tile
    end function

    integer(c_int) function vgpu_set_method(ctx, method) bind(C, name="vgpu_set_method")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: method
    end function

    integer(c_int) function vgpu_assemble_rhs(ctx, q_in, t_old, b) &
        bind(C, name="vgpu_assemble_rhs")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: q_in(*), t_old(*)
      real(c_double), intent(out) :: b(*)
    end function

    integer(c_int) function vgpu_diag(ctx, diag) bind(C, name="vgpu_diag")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: diag(*)
    end function

    integer(c_int) function vgpu_solve(ctx, precision, b, x, rtol, itmax, inner_rtol, &
                                       iters, err) bind(C, name="vgpu_solve")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: precision, itmax
      real(c_double), intent(in) :: b(*)
      real(c_double), intent(inout) :: x(*)
      real(c_double), value :: rtol, inner_rtol
      integer(c_int), intent(out) :: iters
      real(c_double), intent(out) :: err
    end function

    integer(c_int) function vgpu_mass(ctx, c, mass) bind(C, name="vgpu_mass")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: c(*)
      real(c_double), intent(out) :: mass
    end function

    ! diagonal-dominance factor and flag of the current scheme's element
    ! matrices (the host's diagonal_dominance_factor / is_diagonally_dominant)
    integer(c_int) function vgpu_ddf(ctx, ddf, dominant) bind(C, name="vgpu_ddf")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: ddf
      integer(c_int), intent(out) :: dominant
    end function

    ! the host's T-file: x, y, z, c*1e6 per node in e12.5; path NUL-terminated
    integer(c_int) function vgpu_write_tfile(path, n, xx, yy, zz, c) &
        bind(C, name="vgpu_write_tfile")
      import :: c_int, c_int64_t, c_double, c_char
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int64_t), value :: n
      real(c_double), intent(in) :: xx(*), yy(*), zz(*), c(*)
    end function

    ! owner tile shape and z chunk from timings of the solver's iteration launch
    integer(c_int) function vgpu_autotune(ctx, precision, reps, scatter, tile, zchunk, &
                                          best_ms) bind(C, name="vgpu_autotune")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: precision, reps
      integer(c_int), intent(out) :: scatter, tile, zchunk
      real(c_double), intent(out) :: best_ms
    end function

    ! several systems (contexts) solved together; ctxs(nm) holds their c_ptr
    integer(c_int) function vgpu_batch_create(nm, ctxs, batch) &
        bind(C, name="vgpu_batch_create")
      import :: c_int, c_ptr
      integer(c_int), value :: nm
      type(c_ptr), intent(in) :: ctxs(*)
      type(c_ptr), intent(out) :: batch
    end function

    integer(c_int) function vgpu_batch_destroy(batch) bind(C, name="vgpu_batch_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
    end function

    ! b(m), x(m): c_loc of system m's right-hand side and solution
    integer(c_int) function vgpu_batch_solve(batch, precision, b, x, rtol, itmax, iters, &
                                             errs, status) bind(C, name="vgpu_batch_solve")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: batch
      integer(c_int), value :: precision, itmax
      type(c_ptr), intent(in) :: b(*), x(*)
      real(c_double), value :: rtol
      integer(c_int), intent(out) :: iters(*), status(*)
      real(c_double), intent(out) :: errs(*)
    end function

    ! members left at itmax / in breakdown by the batched BiCG are finished by
    ! BiCGSTAB on their own contexts (VGPU_METHOD_BICGSTAB; VGPU_METHOD_BICG: off)
    integer(c_int) function vgpu_batch_set_fallback(batch, method, itmax) &
        bind(C, name="vgpu_batch_set_fallback")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
      integer(c_int), value :: method, itmax
    end function
  end interface
end module vanadis_gpu
