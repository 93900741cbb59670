! giem2g_amd_host.f90 -- Fortran solver driver over libgiem2g_amd.so.
!
!   giem2g_amd_host [deck.json]
!
! Follows the solve part of the reference driver (giem2g.F90:323-507) on one
! GPU: for every frequency of the deck the background is set, the FP64
! Green tensor is built, and the integral equation is solved for the EY and
! then the EX plane-wave polarization (giem2g.F90:346, 435).  The initial
! guess of each solve is a saved solution SOL_P{Y,X}_F<freq>.bin when it
! exists, else the previous frequency's solution of the same polarization
! (none at the first frequency), as in giem2g.F90:351-371.  With
! "Save Solution" the solution is written to that file, complex128 stream
! in (z, component, x, y) order (SaveIESolutionOneFileBinary,
! mpi_saveload_module.F90:568-600).
!
! With "Recalculate Fields" and a "Recievers" list the solution is continued
! to the receivers after each solve (ReCalculation, giem2g.F90:412-428,
! 487-503): the anomalous E and H at every cell centre shifted by each
! receiver's (xshift, yshift) at its depth, plus the normal field, written to
! P{Y,X}_F<freq>.dat in the 28-column text format of SaveOutputOneFile
! (mpi_saveload_module.F90:391-483).  With "Solve Equation": false the
! solutions are read from SOL_P{Y,X}_F<freq>.bin instead of solved, and a
! frequency whose EY solution is missing is skipped (giem2g.F90:390-411).
!
! Deck keys (as the reference's, giem2g.F90 and mpi_saveload_module.F90):
!   Background.Conductivity [Nl], Background.Thickness [Nl-1],
!   Anomaly.{Nx, Ny, Nz, dx, dy, zborders[Nz+1], Conductivity (file name)},
!   Frequencies [...], FGMRES.{OuterBufferSize, InnerBufferSize, MaxRestart,
!   Tolerance}, "Save Solution", "Solve Equation", "Recalculate Fields",
!   Recievers [{xshift, yshift, depth}].
! The conductivity file holds Nx*Ny*Nz values in (Nz, Nx, Ny) order
! (LoadAnomalySigma, mpi_saveload_module.F90:113-147).
! Keys of this build: Solver.{InnerPrecision (0 FP64, 1 complex64),
! Precond (0, 1, 2), Network (FP32 parameter file), Accept}.
program giem2g_amd_host
  use iso_c_binding
  use g2g_json
  use giem2g_amd_api
  implicit none

  double precision, parameter :: PI = 3.14159265358979323846264338327950288d0
  double precision, parameter :: MU0 = 4d-7 * PI
  double precision, parameter :: EPS0 = 1d0 / (MU0 * 299792458d0**2)
  double precision, parameter :: AIR_SIGMA = 1d-12

  type(json_doc) :: deck
  character(len=4096) :: deck_path, sig_path, net_path, fname
  character(len=22) :: fnum
  logical :: ok, found, save_solution, have_guess, solve_equation, recalc
  integer :: nl, nx, ny, nz, nf, n, i, ix, iy, iz, ic, ifreq, ipol, pol, rc, ncell, u, ios, nr, ir
  double precision :: dx, dy, omega, accept
  double precision, allocatable :: sigb(:), thick(:), z(:), freqs(:), siga(:, :, :)
  complex(c_double_complex), allocatable :: csigb(:), csiga(:), ebkg(:), esol(:)
  complex(c_double_complex), allocatable, target :: eprev(:, :)
  complex(c_double_complex), allocatable :: e4(:, :, :, :), f4(:, :, :, :)
  real(c_float), allocatable :: wnet(:)
  integer(c_long_long) :: np
  type(c_ptr) :: op, net, rcop
  double precision, allocatable :: rxs(:), rys(:), rdepth(:), rzr(:)
  integer(c_int), allocatable :: rlayer(:), rcell(:)
  complex(c_double_complex), allocatable :: ea(:, :, :, :), ha(:, :, :, :), f0(:, :)
  type(giem2g_solver_params) :: prm
  type(giem2g_solve_info) :: info
  logical :: have_prev(0:1)
  character(len=2), parameter :: pname(0:1) = ['PX', 'PY']
  character(len=4096) :: arg
  logical :: check_only

  deck_path = 'giem2g_input.json'
  check_only = .false.
  do i = 1, command_argument_count()
     call get_command_argument(i, arg)
     if (trim(arg) == '--check') then
        check_only = .true.
     else
        deck_path = arg
     end if
  end do
  call json_load(trim(deck_path), deck, ok)
  if (.not. ok) call fail('cannot read deck ' // trim(deck_path))

  nl = json_array_len(deck, 'Background.Conductivity')
  if (nl < 1) call fail('Background.Conductivity missing')
  allocate(sigb(nl), thick(max(nl - 1, 1)), csigb(0:nl))
  call json_get_real_array(deck, 'Background.Conductivity', sigb, n)
  thick = 0d0
  if (nl > 1) then
     call json_get_real_array(deck, 'Background.Thickness', thick, n)
     if (n /= nl - 1) call fail('Background.Thickness must have Nl-1 entries')
  end if
  nx = json_get_int(deck, 'Anomaly.Nx', 0)
  ny = json_get_int(deck, 'Anomaly.Ny', 0)
  nz = json_get_int(deck, 'Anomaly.Nz', 0)
  dx = json_get_real(deck, 'Anomaly.dx', 0d0)
  dy = json_get_real(deck, 'Anomaly.dy', 0d0)
  allocate(z(nz + 1))
  call json_get_real_array(deck, 'Anomaly.zborders', z, n)
  if (n /= nz + 1) call fail('Anomaly.zborders must have Nz+1 entries')
  nf = json_array_len(deck, 'Frequencies')
  if (nf < 1) call fail('Frequencies missing')
  allocate(freqs(nf))
  call json_get_real_array(deck, 'Frequencies', freqs, n)
  prm%outer_m = json_get_int(deck, 'FGMRES.OuterBufferSize', 30)
  prm%inner_m = json_get_int(deck, 'FGMRES.InnerBufferSize', 30)
  prm%max_restart = json_get_int(deck, 'FGMRES.MaxRestart', 100)
  prm%tol = json_get_real(deck, 'FGMRES.Tolerance', 1d-8)
  prm%inner_precision = json_get_int(deck, 'Solver.InnerPrecision', 0)
  prm%precond = json_get_int(deck, 'Solver.Precond', 0)
  prm%verbose = 0
  prm%learned_patience = 0
  prm%inner_tol = json_get_real(deck, 'Solver.InnerTolerance', 0d0)
  accept = json_get_real(deck, 'Solver.Accept', 1d0)
  save_solution = json_get_logical(deck, 'Save Solution', .true.)
  solve_equation = json_get_logical(deck, 'Solve Equation', .true.)
  ! receivers (LoadRecievers, mpi_saveload_module.F90:174-213)
  recalc = json_get_logical(deck, 'Recalculate Fields', .true.)
  nr = max(json_array_len(deck, 'Recievers'), 0)
  if (recalc .and. nr == 0) then
     write (*, '(A)') 'no Recievers in the deck: fields are not continued to receivers'
     recalc = .false.
  end if
  if (recalc) then
     allocate(rxs(nr), rys(nr), rdepth(nr), rzr(nr), rlayer(nr), rcell(nr))
     do ir = 1, nr
        rxs(ir) = json_get_elem_real(deck, 'Recievers', ir, 'xshift', 0d0)
        rys(ir) = json_get_elem_real(deck, 'Recievers', ir, 'yshift', 0d0)
        rdepth(ir) = json_get_elem_real(deck, 'Recievers', ir, 'depth', 0d0)
     end do
  end if

  ! anomaly conductivity, file order (Nz, Nx, Ny)
  call json_get_string(deck, 'Anomaly.Conductivity', sig_path, found)
  if (.not. found) call fail('Anomaly.Conductivity (file name) missing')
  if (check_only) then
     ! --check: print the parsed deck and stop before any GPU work
     write (*, '(A,I0)') 'Nl ', nl
     write (*, '(A,*(ES15.7))') 'Conductivity ', sigb
     if (nl > 1) write (*, '(A,*(ES15.7))') 'Thickness ', thick(1:nl - 1)
     write (*, '(A,3(1X,I0),2(1X,ES15.7))') 'Grid', nx, ny, nz, dx, dy
     write (*, '(A,ES15.7,1X,ES15.7)') 'zborders ', z(1), z(nz + 1)
     write (*, '(A,*(ES15.7))') 'Frequencies ', freqs
     write (*, '(A,3(1X,I0),1X,ES15.7)') 'FGMRES', prm%outer_m, prm%inner_m, prm%max_restart, prm%tol
     write (*, '(A,L1)') 'SaveSolution ', save_solution
     write (*, '(A,L1)') 'SolveEquation ', solve_equation
     write (*, '(A,L1)') 'RecalculateFields ', recalc
     do ir = 1, merge(nr, 0, recalc)
        write (*, '(A,3(1X,ES15.7))') 'Reciever', rxs(ir), rys(ir), rdepth(ir)
     end do
     write (*, '(A,A)') 'SigmaFile ', trim(sig_path)
     stop
  end if
  allocate(siga(nz, nx, ny))
  open(newunit=u, file=trim(sig_path), status='old', iostat=ios)
  if (ios /= 0) call fail('cannot open conductivity file ' // trim(sig_path))
  read(u, *, iostat=ios) siga
  close(u)
  if (ios /= 0) call fail('conductivity file too short')

  ncell = nx * ny * nz
  allocate(csiga(ncell), ebkg(3 * ncell), esol(3 * ncell), eprev(3 * ncell, 0:1))
  allocate(e4(nx, ny, nz, 3), f4(nz, 3, nx, ny))
  have_prev = .false.

  rc = giem2g_op_create(nx, ny, nz, dx, dy, z, op)
  call check(rc, 'giem2g_op_create')
  rcop = c_null_ptr
  if (recalc) then
     call check(giem2g_rc_create(op, nr, rxs, rys, rdepth, rcop), 'giem2g_rc_create')
     allocate(ea(nr, 3, nx, ny), ha(nr, 3, nx, ny), f0(nr, 6))
  end if
  net = c_null_ptr
  if (prm%precond /= 0) then
     call json_get_string(deck, 'Solver.Network', net_path, found)
     if (.not. found) call fail('Solver.Precond needs Solver.Network')
     call check(giem2g_unet_nparams(np), 'giem2g_unet_nparams')
     allocate(wnet(np))
     open(newunit=u, file=trim(net_path), access='stream', form='unformatted', status='old', iostat=ios)
     if (ios /= 0) call fail('cannot open network file ' // trim(net_path))
     read(u, iostat=ios) wnet
     close(u)
     if (ios /= 0) call fail('network file too short')
     call check(giem2g_unet_create(net), 'giem2g_unet_create')
     call check(giem2g_unet_set_params(net, wnet), 'giem2g_unet_set_params')
     call check(giem2g_op_set_learned(op, net, accept), 'giem2g_op_set_learned')
  end if

  do ifreq = 1, nf
     omega = 2d0 * PI * freqs(ifreq)
     ! SET_FREQ (data_types_module.F90:105-120): sigma - i omega eps0, air first
     csigb(0) = cmplx(AIR_SIGMA, -omega * EPS0, c_double_complex)
     do i = 1, nl
        csigb(i) = cmplx(sigb(i), -omega * EPS0, c_double_complex)
     end do
     call check(giem2g_op_set_background(op, nl, csigb, thick, omega), 'giem2g_op_set_background')
     if (solve_equation) call check(giem2g_op_build_tensor(op), 'giem2g_op_build_tensor')
     ! SetAnomalySigma (ie_solver_module.F90:129-131): (Nz,Nx,Ny) -> (Nx,Ny,Nz)
     do iy = 1, ny
        do ix = 1, nx
           do iz = 1, nz
              csiga(ix + nx * (iy - 1) + nx * ny * (iz - 1)) = &
                   cmplx(siga(iz, ix, iy), -omega * EPS0, c_double_complex)
           end do
        end do
     end do
     if (solve_equation) call check(giem2g_op_set_sigma(op, csiga), 'giem2g_op_set_sigma')
     if (recalc) then
        ! CalcRecalculationGreenTensor, CalcFFTofRCTensor (giem2g.F90:340-343),
        ! FinalizeRCOperator
        call check(giem2g_rc_calc_kernel(rcop), 'giem2g_rc_calc_kernel')
        call check(giem2g_rc_calc_fft_of_kernel(rcop), 'giem2g_rc_calc_fft_of_kernel')
        call check(giem2g_rc_set_anomaly_conductivity(rcop, csiga), 'giem2g_rc_set_anomaly_conductivity')
        call check(giem2g_rc_receivers(rcop, rzr, rlayer, rcell), 'giem2g_rc_receivers')
     end if
     write (fnum, '(F22.10)') freqs(ifreq)
     do i = 1, 11
        if (fnum(i:i) == ' ') fnum(i:i) = '0'
     end do
     do ipol = 1, 2
        pol = GIEM2G_PW_EY
        if (ipol == 2) pol = GIEM2G_PW_EX
        fname = 'SOL_' // pname(pol) // '_F' // trim(fnum) // '.bin'
        inquire (file=trim(fname), exist=have_guess)
        if (have_guess) then
           open(newunit=u, file=trim(fname), access='stream', form='unformatted', status='old')
           read(u) f4
           close(u)
           do ic = 1, 3
              do iz = 1, nz
                 e4(:, :, iz, ic) = f4(iz, ic, :, :)
              end do
           end do
           eprev(:, pol) = reshape(e4, [3 * ncell])
           have_prev(pol) = .true.
           write (*, '(A)') 'Load IE Solution from ' // trim(fname)
        end if
        if (.not. solve_equation) then
           ! recalculation only (giem2g.F90:390-411, 473-486)
           if (.not. have_guess) then
              write (*, '(A)') 'Solution not found: ' // trim(fname)
              exit
           end if
           call recalculate(eprev(:, pol))
           cycle
        end if
        call check(giem2g_plane_wave(op, pol, ebkg), 'giem2g_plane_wave')
        if (have_prev(pol)) then
           call check(giem2g_solve(op, ebkg, c_loc(eprev(1, pol)), prm, esol, info), 'giem2g_solve')
        else
           call check(giem2g_solve(op, ebkg, c_null_ptr, prm, esol, info), 'giem2g_solve')
        end if
        eprev(:, pol) = esol
        have_prev(pol) = .true.
        write (*, '(A,ES12.5,A,A,A,I4,A,I6,A,I6,A,ES10.3,A,I3,A,F9.4,A)') 'freq ', freqs(ifreq), &
             ' Hz  ', pname(pol), '  outer ', info%outer_iterations, '  mv64 ', info%matvecs_fp64, &
             '  mv32 ', info%matvecs_fp32, '  true residual ', info%true_residual, '  status ', &
             info%status, '  ', info%seconds, ' s'
        if (save_solution) then
           e4 = reshape(esol, [nx, ny, nz, 3])
           do ic = 1, 3
              do iz = 1, nz
                 f4(iz, ic, :, :) = e4(:, :, iz, ic)
              end do
           end do
           open(newunit=u, file=trim(fname), access='stream', form='unformatted', status='replace')
           write (u) f4
           close(u)
        end if
        if (recalc) call recalculate(esol)
     end do
  end do
  if (c_associated(rcop)) call check(giem2g_rc_destroy(rcop), 'giem2g_rc_destroy')
  if (c_associated(net)) then
     call check(giem2g_op_set_learned(op, c_null_ptr, 1d0), 'giem2g_op_set_learned')
     call check(giem2g_unet_destroy(net), 'giem2g_unet_destroy')
  end if
  call check(giem2g_op_destroy(op), 'giem2g_op_destroy')

contains

  ! ReCalculation of the field E of polarization pol at the receivers, the
  ! total field with PlaneWave (giem2g.F90:412-426), and SaveOutputOneFile
  ! (mpi_saveload_module.F90:391-483) to P{Y,X}_F<freq>.dat
  subroutine recalculate(e)
    complex(c_double_complex), intent(in) :: e(:)
    character(len=*), parameter :: title_fmt = '(A1, 28A20)', output_fmt = '(A1, 28ES20.10E3)'
    character(len=4096) :: oname
    double precision :: x, y
    integer :: uo, k, jr, jx, jy
    call check(giem2g_rc_recalculate(rcop, e, ea, ha), 'giem2g_rc_recalculate')
    call check(giem2g_rc_plane_wave(rcop, pol, f0), 'giem2g_rc_plane_wave')
    oname = pname(pol) // '_F' // trim(fnum) // '.dat'
    open(newunit=uo, file=trim(oname), status='replace')
    write (uo, title_fmt) '%', 'x_recv', 'y_recv', 'z_recv', 'frequency', &
         'Re Ex^a', 'Im Ex^a', 'Re Ey^a', 'Im Ey^a', 'Re Ez^a', 'Im Ez^a', &
         'Re Hx^a', 'Im Hx^a', 'Re Hy^a', 'Im Hy^a', 'Re Hz^a', 'Im Hz^a', &
         'Re Ex^t', 'Im Ex^t', 'Re Ey^t', 'Im Ey^t', 'Re Ez^t', 'Im Ez^t', &
         'Re Hx^t', 'Im Hx^t', 'Re Hy^t', 'Im Hy^t', 'Re Hz^t', 'Im Hz^t'
    do jr = 1, nr
       do jy = 1, ny
          y = jy * dy - dy / 2d0 + rys(jr)
          do jx = 1, nx
             x = jx * dx - dx / 2d0 + rxs(jr)
             write (uo, output_fmt) ' ', x, y, rzr(jr), freqs(ifreq), &
                  (ea(jr, k, jx, jy), k = 1, 3), (ha(jr, k, jx, jy), k = 1, 3), &
                  (ea(jr, k, jx, jy) + f0(jr, k), k = 1, 3), (ha(jr, k, jx, jy) + f0(jr, 3 + k), k = 1, 3)
          end do
       end do
    end do
    close(uo)
    write (*, '(A)') 'Fields at receivers written to ' // trim(oname)
  end subroutine recalculate

  subroutine check(code, what)
    integer(c_int), intent(in) :: code
    character(len=*), intent(in) :: what
    if (code /= 0) call fail(what // ': ' // last_error())
  end subroutine check

  subroutine fail(msg)
    character(len=*), intent(in) :: msg
    write (*, '(A)') 'giem2g_amd_host: ' // msg
    error stop 1
  end subroutine fail

end program giem2g_amd_host
