! giem2g_amd_api.f90 -- ISO_C_BINDING interfaces of libgiem2g_amd.so
! (include/giem2g_amd.h) for Fortran hosts.  The nine reference entry
! points (giem2g_interfaces.F90:76-254) keep the reference's own BIND(C)
! interfaces; this module binds the status-returning operator, solver and
! learned-preconditioner entry points that a Fortran driver needs.
module giem2g_amd_api
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: GIEM2G_PW_EX = 0, GIEM2G_PW_EY = 1

  type, bind(c) :: giem2g_solver_params
     integer(c_int) :: outer_m, inner_m, max_restart
     real(c_double) :: tol
     integer(c_int) :: inner_precision, precond, verbose, learned_patience
     real(c_double) :: inner_tol
  end type giem2g_solver_params

  type, bind(c) :: giem2g_solve_info
     integer(c_int) :: outer_iterations, status
     real(c_double) :: arnoldi_residual, true_residual
     integer(c_long_long) :: matvecs_fp64, matvecs_fp32, precond_calls, precond_fallbacks
     real(c_double) :: seconds
  end type giem2g_solve_info

  interface
     function giem2g_last_error() bind(c) result(p)
       import :: c_ptr
       type(c_ptr) :: p
     end function giem2g_last_error

     integer(c_int) function giem2g_op_create(nx, ny, nz, dx, dy, z, op) bind(c)
       import :: c_int, c_double, c_ptr
       integer(c_int), value :: nx, ny, nz
       real(c_double), value :: dx, dy
       real(c_double), intent(in) :: z(*)
       type(c_ptr), intent(out) :: op
     end function giem2g_op_create

     integer(c_int) function giem2g_op_destroy(op) bind(c)
       import :: c_int, c_ptr
       type(c_ptr), value :: op
     end function giem2g_op_destroy

     integer(c_int) function giem2g_op_set_background(op, nl, csigb, thick, omega) bind(c)
       import :: c_int, c_double, c_double_complex, c_ptr
       type(c_ptr), value :: op
       integer(c_int), value :: nl
       complex(c_double_complex), intent(in) :: csigb(*)
       real(c_double), intent(in) :: thick(*)
       real(c_double), value :: omega
     end function giem2g_op_set_background

     integer(c_int) function giem2g_op_build_tensor(op) bind(c)
       import :: c_int, c_ptr
       type(c_ptr), value :: op
     end function giem2g_op_build_tensor

     integer(c_int) function giem2g_op_set_sigma(op, csiga) bind(c)
       import :: c_int, c_double_complex, c_ptr
       type(c_ptr), value :: op
       complex(c_double_complex), intent(in) :: csiga(*)
     end function giem2g_op_set_sigma

     integer(c_int) function giem2g_plane_wave(op, pol, e) bind(c)
       import :: c_int, c_double_complex, c_ptr
       type(c_ptr), value :: op
       integer(c_int), value :: pol
       complex(c_double_complex), intent(out) :: e(*)
     end function giem2g_plane_wave

     integer(c_int) function giem2g_solve(op, ebkg, e0, p, esol, info) bind(c)
       import :: c_int, c_double_complex, c_ptr, giem2g_solver_params, giem2g_solve_info
       type(c_ptr), value :: op
       complex(c_double_complex), intent(in) :: ebkg(*)
       type(c_ptr), value :: e0
       type(giem2g_solver_params), intent(in) :: p
       complex(c_double_complex), intent(out) :: esol(*)
       type(giem2g_solve_info), intent(out) :: info
     end function giem2g_solve

     integer(c_int) function giem2g_unet_create(net) bind(c)
       import :: c_int, c_ptr
       type(c_ptr), intent(out) :: net
     end function giem2g_unet_create

     integer(c_int) function giem2g_unet_destroy(net) bind(c)
       import :: c_int, c_ptr
       type(c_ptr), value :: net
     end function giem2g_unet_destroy

     integer(c_int) function giem2g_unet_nparams(n) bind(c)
       import :: c_int, c_long_long
       integer(c_long_long), intent(out) :: n
     end function giem2g_unet_nparams

     integer(c_int) function giem2g_unet_set_params(net, p) bind(c)
       import :: c_int, c_float, c_ptr
       type(c_ptr), value :: net
       real(c_float), intent(in) :: p(*)
     end function giem2g_unet_set_params

     integer(c_int) function giem2g_op_set_learned(op, net, accept) bind(c)
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: op, net
       real(c_double), value :: accept
     end function giem2g_op_set_learned
     ! continuation of the solution to receivers
     integer(c_int) function giem2g_rc_create(ie_op, nr, xshift, yshift, depth, rc) bind(c)
       import
       type(c_ptr), value :: ie_op
       integer(c_int), value :: nr
       real(c_double) :: xshift(*), yshift(*), depth(*)
       type(c_ptr) :: rc
     end function giem2g_rc_create

     integer(c_int) function giem2g_rc_destroy(rc) bind(c)
       import
       type(c_ptr), value :: rc
     end function giem2g_rc_destroy

     integer(c_int) function giem2g_rc_calc_kernel(rc) bind(c)
       import
       type(c_ptr), value :: rc
     end function giem2g_rc_calc_kernel

     integer(c_int) function giem2g_rc_calc_fft_of_kernel(rc) bind(c)
       import
       type(c_ptr), value :: rc
     end function giem2g_rc_calc_fft_of_kernel

     integer(c_int) function giem2g_rc_receivers(rc, zr, layer, cell) bind(c)
       import
       type(c_ptr), value :: rc
       real(c_double) :: zr(*)
       integer(c_int) :: layer(*), cell(*)
     end function giem2g_rc_receivers

     integer(c_int) function giem2g_rc_set_anomaly_conductivity(rc, csiga) bind(c)
       import
       type(c_ptr), value :: rc
       complex(c_double_complex) :: csiga(*)
     end function giem2g_rc_set_anomaly_conductivity

     integer(c_int) function giem2g_rc_recalculate(rc, eint, ea, ha) bind(c)
       import
       type(c_ptr), value :: rc
       complex(c_double_complex) :: eint(*), ea(*), ha(*)
     end function giem2g_rc_recalculate

     integer(c_int) function giem2g_rc_plane_wave(rc, pol, f0) bind(c)
       import
       type(c_ptr), value :: rc
       integer(c_int), value :: pol
       complex(c_double_complex) :: f0(*)
     end function giem2g_rc_plane_wave
  end interface

contains

  function last_error() result(msg)
    ! message of the last failed call as a Fortran string
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: s(:)
    type(c_ptr) :: p
    integer :: n
    p = giem2g_last_error()
    msg = ''
    if (.not. c_associated(p)) return
    call c_f_pointer(p, s, [4096])
    n = 0
    do while (n < 4096)
       if (s(n + 1) == c_null_char) exit
       n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function last_error

end module giem2g_amd_api
