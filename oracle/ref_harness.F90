! ref_harness.F90 -- test harness around the reference's own Fortran modules.
!
! TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
! the reference sources where they lie under /root/reference (Hankel filter,
! layered kernel, GFGMRES) into oracle/_ref/liboracle_ref.so.  The tests use
! it to check the C oracle against the reference code itself.  Nothing here
! travels into the product.
MODULE REF_HARNESS
  USE, INTRINSIC :: ISO_C_BINDING
  USE CONST_MODULE
  USE DATA_TYPES_MODULE
  USE APQ_MODULE
  USE IE_Kernel_Image_Module
  USE RC_Kernel_Image_Module, ONLY: Calc_Integral_U
  USE VolumeBesselTransforms
  USE IntegralCodes
  USE FGMRES_INTERFACES
  USE SEQ_FGMRES
  IMPLICIT NONE
  COMPLEX(REALPARM), POINTER :: dense_A(:,:) => NULL()
  INTEGER :: dense_N = 0
  INTEGER(8) :: n_matvec = 0
CONTAINS

  SUBROUTINE make_bkg(Nl, csig, thick, omega, bkg)
    INTEGER(C_INT), INTENT(IN) :: Nl
    REAL(C_DOUBLE), INTENT(IN) :: csig(2*(Nl+1)), thick(*), omega
    TYPE(BKG_DATA_TYPE), INTENT(INOUT) :: bkg
    INTEGER :: I
    bkg%Nl = Nl
    ALLOCATE(bkg%csigma(0:Nl), bkg%k(0:Nl), bkg%k2(0:Nl))
    ALLOCATE(bkg%thick(1:MAX(Nl-1,1)), bkg%depth(1:MAX(Nl-1,1)))
    DO I = 0, Nl
      bkg%csigma(I) = CMPLX(csig(2*I+1), csig(2*I+2), KIND=REALPARM)
    ENDDO
    IF (Nl > 1) THEN
      bkg%thick(1:Nl-1) = thick(1:Nl-1)
      bkg%depth(1) = bkg%thick(1)
      DO I = 2, Nl-1
        bkg%depth(I) = bkg%depth(I-1) + bkg%thick(I)
      ENDDO
    ENDIF
    bkg%omega = omega
    bkg%freq = omega/PI/2
    bkg%k2 = C_IONE*bkg%csigma*MU0*bkg%omega
    bkg%k = SQRT(bkg%k2)
    bkg%iwm = C_IONE*MU0*bkg%omega
  END SUBROUTINE

  SUBROUTINE make_anomaly(Nz, z, bkg, anomaly)
    INTEGER(C_INT), INTENT(IN) :: Nz
    REAL(C_DOUBLE), INTENT(IN) :: z(Nz+1)
    TYPE(BKG_DATA_TYPE), INTENT(IN) :: bkg
    TYPE(ANOMALY_TYPE), INTENT(INOUT) :: anomaly
    anomaly%Nz = Nz
    ALLOCATE(anomaly%z(0:Nz), anomaly%Lnumber(0:Nz), anomaly%dz(Nz))
    anomaly%z(0:Nz) = z(1:Nz+1)
    anomaly%Lnumber(1:Nz) = GetLayer((anomaly%z(0:Nz-1)+anomaly%z(1:Nz))/2d0, bkg)
    anomaly%Lnumber(0) = GetLayer(anomaly%z(0)*(1d0-1d-7), bkg)
    anomaly%dz = anomaly%z(1:Nz) - anomaly%z(0:Nz-1)
  END SUBROUTINE

  SUBROUTINE ref_vbt_weights(rmin, x, y, hx, hy, WT, lms) BIND(C, NAME='ref_vbt_weights')
    REAL(C_DOUBLE), VALUE, INTENT(IN) :: rmin, x, y, hx, hy
    REAL(C_DOUBLE), INTENT(OUT) :: WT(Nfirst:Nlast, 6), lms(Nfirst:Nlast)
    CALL VBTransformInit(rmin, lms)
    CALL VBTransformWeightsAllDoubleInt(x, y, hx, hy, WT)
  END SUBROUTINE

  SUBROUTINE ref_vbt_weights_single(rmin, x, y, hx, hy, WT, lms) BIND(C, NAME='ref_vbt_weights_single')
    REAL(C_DOUBLE), VALUE, INTENT(IN) :: rmin, x, y, hx, hy
    REAL(C_DOUBLE), INTENT(OUT) :: WT(Nfirst:Nlast, 6), lms(Nfirst:Nlast)
    CALL VBTransformInit(rmin, lms)
    CALL VBTransformWeightsAllSingleInt(x, y, hx, hy, WT)
  END SUBROUTINE

  ! PrepareRecvs on Nr receivers at the given depths
  SUBROUTINE ref_prepare_recvs(Nl, csig, thick, omega, Nz, z, Nr, depth, zr, layer, cell) &
      BIND(C, NAME='ref_prepare_recvs')
    INTEGER(C_INT), VALUE, INTENT(IN) :: Nl, Nz, Nr
    REAL(C_DOUBLE), INTENT(IN) :: csig(2*(Nl+1)), thick(*), z(Nz+1), depth(Nr)
    REAL(C_DOUBLE), VALUE, INTENT(IN) :: omega
    REAL(C_DOUBLE), INTENT(OUT) :: zr(Nr)
    INTEGER(C_INT), INTENT(OUT) :: layer(Nr), cell(Nr)
    TYPE(BKG_DATA_TYPE) :: bkg
    TYPE(ANOMALY_TYPE) :: anomaly
    TYPE(RECEIVER_TYPE), POINTER :: recvs(:)
    INTEGER :: I
    CALL make_bkg(Nl, csig, thick, omega, bkg)
    CALL make_anomaly(Nz, z, bkg, anomaly)
    ALLOCATE(recvs(Nr))
    DO I = 1, Nr
      recvs(I)%zrecv = depth(I)
      recvs(I)%x_shift = 0
      recvs(I)%y_shift = 0
    ENDDO
    CALL PrepareRecvs(recvs, anomaly, bkg)
    DO I = 1, Nr
      zr(I) = recvs(I)%zrecv
      layer(I) = recvs(I)%recv_layer
      cell(I) = recvs(I)%anom_cell
    ENDDO
    DEALLOCATE(recvs)
  END SUBROUTINE

  ! Calc_Integral_U of the receiver continuation for one receiver (depth as
  ! given, prepared by PrepareRecvs) and one lambda
  SUBROUTINE ref_rc_integral_u(Nl, csig, thick, omega, Nz, z, depth, lm, W, GE, GH) &
      BIND(C, NAME='ref_rc_integral_u')
    INTEGER(C_INT), VALUE, INTENT(IN) :: Nl, Nz
    REAL(C_DOUBLE), INTENT(IN) :: csig(2*(Nl+1)), thick(*), z(Nz+1), W(6)
    REAL(C_DOUBLE), VALUE, INTENT(IN) :: omega, depth, lm
    COMPLEX(C_DOUBLE_COMPLEX), INTENT(INOUT) :: GE(8, Nz), GH(7, Nz)
    TYPE(BKG_DATA_TYPE) :: bkg
    TYPE(ANOMALY_TYPE) :: anomaly
    TYPE(RECEIVER_TYPE), POINTER :: recvs(:)
    REAL(REALPARM) :: ctime(2)
    CALL make_bkg(Nl, csig, thick, omega, bkg)
    CALL make_anomaly(Nz, z, bkg, anomaly)
    ALLOCATE(recvs(1))
    recvs(1)%zrecv = depth
    recvs(1)%x_shift = 0
    recvs(1)%y_shift = 0
    CALL PrepareRecvs(recvs, anomaly, bkg)
    ctime = 0
    CALL Calc_Integral_U(anomaly, bkg, recvs(1), lm, W, GE, GH, ctime)
    DEALLOCATE(recvs)
  END SUBROUTINE

  SUBROUTINE ref_double_integral_u(Nl, csig, thick, omega, Nz, z, lm, W, G) &
      BIND(C, NAME='ref_double_integral_u')
    INTEGER(C_INT), VALUE, INTENT(IN) :: Nl, Nz
    REAL(C_DOUBLE), INTENT(IN) :: csig(2*(Nl+1)), thick(*), z(Nz+1), W(6)
    REAL(C_DOUBLE), VALUE, INTENT(IN) :: omega, lm
    COMPLEX(C_DOUBLE_COMPLEX), INTENT(INOUT) :: G(6, Nz, Nz)
    TYPE(BKG_DATA_TYPE) :: bkg
    TYPE(ANOMALY_TYPE) :: anomaly
    REAL(REALPARM) :: ctime(2)
    CALL make_bkg(Nl, csig, thick, omega, bkg)
    CALL make_anomaly(Nz, z, bkg, anomaly)
    ctime = 0
    CALL Calc_Double_Integral_U(anomaly, bkg, anomaly%dz, lm, W, G, ctime)
  END SUBROUTINE

  SUBROUTINE ref_calc_apq(Nl, csig, thick, omega, lm, Arr, p, q, eta, A) &
      BIND(C, NAME='ref_calc_apq')
    INTEGER(C_INT), VALUE, INTENT(IN) :: Nl
    REAL(C_DOUBLE), INTENT(IN) :: csig(2*(Nl+1)), thick(*)
    REAL(C_DOUBLE), VALUE, INTENT(IN) :: omega, lm
    COMPLEX(C_DOUBLE_COMPLEX), INTENT(OUT) :: Arr(Nl,2), p(Nl,2), q(Nl,2), eta(Nl), A(Nl,Nl,2)
    TYPE(BKG_DATA_TYPE) :: bkg
    CALL make_bkg(Nl, csig, thick, omega, bkg)
    A = C_ZERO
    CALL Calc_APQ(bkg, lm, Arr, p, q, eta, A)
  END SUBROUTINE

  ! ---- GFGMRES on a dense complex matrix ----
  SUBROUTINE dense_apply(pA, v_in, v_out)
    TYPE(C_PTR), INTENT(IN) :: pA
    COMPLEX(REALPARM), INTENT(IN) :: v_in(:)
    COMPLEX(REALPARM), INTENT(INOUT) :: v_out(:)
    v_out = MATMUL(dense_A, v_in)
    n_matvec = n_matvec + 1
  END SUBROUTINE

  SUBROUTINE dense_dp(M, v, K, res, ptr)
    COMPLEX(REALPARM), INTENT(IN) :: M(:,:)
    COMPLEX(REALPARM), INTENT(IN) :: v(:)
    INTEGER, INTENT(IN) :: K
    COMPLEX(REALPARM), INTENT(INOUT) :: res(:)
    TYPE(C_PTR), INTENT(IN) :: ptr
    INTEGER :: N
    N = SIZE(v)
    CALL ZGEMV('C', N, K, C_ONE, M, N, v, 1, C_ZERO, res, 1)
  END SUBROUTINE

  FUNCTION quiet_inform(info) RESULT(interp)
    TYPE(RESULT_INFO), INTENT(IN) :: info
    INTEGER :: interp
    interp = NOINTERUPT
  END FUNCTION

  SUBROUTINE ref_fgmres_dense(N, Amat, b, x0, m_out, m_in, maxit, tol, x, iters, stat, bea, be, nmv) &
      BIND(C, NAME='ref_fgmres_dense')
    INTEGER(C_INT), VALUE, INTENT(IN) :: N, m_out, m_in, maxit
    REAL(C_DOUBLE), VALUE, INTENT(IN) :: tol
    COMPLEX(C_DOUBLE_COMPLEX), TARGET, INTENT(IN) :: Amat(N, N), b(N), x0(N)
    COMPLEX(C_DOUBLE_COMPLEX), INTENT(OUT) :: x(N)
    INTEGER(C_INT), INTENT(OUT) :: iters, stat
    REAL(C_DOUBLE), INTENT(OUT) :: bea, be
    INTEGER(C_LONG_LONG), INTENT(OUT) :: nmv
    TYPE(SEQ_FGMRES_DATA) :: seq_solver
    PROCEDURE(MatrixVectorMult), POINTER :: Apply
    PROCEDURE(MANYDOTPRODUCTS), POINTER :: MANYDP
    PROCEDURE(InformAboutIteration), POINTER :: InformMe
    TYPE(GMRES_PARAMETERS) :: params
    TYPE(RESULT_INFO) :: info
    INTEGER :: MM(2)
    INTEGER(C_INTPTR_T) :: length
    COMPLEX(REALPARM), ALLOCATABLE :: sol(:), guess(:), rhs(:)
    Apply => dense_apply
    MANYDP => dense_dp
    InformMe => quiet_inform
    IF (ASSOCIATED(dense_A)) DEALLOCATE(dense_A)
    ALLOCATE(dense_A(N, N))
    dense_A = Amat
    dense_N = N
    n_matvec = 0
    params%MaxIt = maxit
    params%Tol = tol
    params%RESTART_RESIDUAL = .FALSE.
    MM(1) = m_out
    MM(2) = m_in
    CALL INIT_SEQ_FGMRES(seq_solver, N, MM, 2, params, length)
    CALL SET_OPERATORS(seq_solver, Apply, MANYDP, InformMe, C_NULL_PTR)
    CALL SET_DP_INSTANCE(seq_solver, C_NULL_PTR)
    ALLOCATE(sol(N), guess(N), rhs(N))
    sol = C_ONE
    guess = x0
    rhs = b
    CALL SEQ_FGMRES_SOLVE(seq_solver, sol, guess, rhs, info)
    x = sol
    iters = info%Iterations
    stat = info%stat
    bea = info%bea
    be = info%be
    nmv = n_matvec
    CALL DELETE_SEQ_FGMRES(seq_solver)
    DEALLOCATE(sol, guess, rhs)
  END SUBROUTINE
END MODULE
