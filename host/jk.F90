! jk.F90 -- the drop-in programs djk.exe (real) and zjk.exe (built with -DZJK):
!
!   djk.exe FN N [N_2 [ID_NCP [ID_TRU]]]
!
! Same command line, files and messages as the reference's djk.exe/zjk.exe
! (src/D/djk.F90, src/Z/zjk.F90, src/shared/readcl.F90): FN.Y (N x N, FP64 or
! complex128, column-major) and FN.J (N int64 signs) in; FN.YU, FN.ZZ (N x N)
! and FN.SY (N FP64) out; two stdout lines; the per-step CSV on stderr, line by
! line while the loop runs (the lines of each batch of steps the host polls
! for together).  The HSVD itself runs on the GPU through libjkg.so.
!
! The solver (environment variable JKG_SOLVER):
!   plain   (default) jkg_dsolve / jkg_zsolve: the reference's arithmetic and
!           trajectory, bit for bit (YU, ZZ, SY and every CSV field but the
!           two timing columns)
!   mixed   jkg_dsolve_mixed / jkg_zsolve_mixed: Gram-form hyperbolic weights
!           on the matrix cores and the relative convergence test
!   learned the mixed solver with the trained step controller, whose
!           JKG_CTL_NPARAM FP32 parameters are read from the file named by
!           JKG_CTL_PARAMS (default: jkogb_amd/ctl_params.f32 next to this
!           program's directory) -- real problems only
!   warm    the mixed solver starting with FP32 steps (fp32_warm)
! Other environment: JKG_DEVICE (default 0), JKG_FUSED=1 selects FMAD = fma.
!
! Deviations from the reference: a failure before the loop (arguments, files,
! the device) ends with ERROR STOP (nonzero exit status; the reference's STOP
! exits with 0) and the reference's message; BUILDs and TRANSFs are device
! times; ID_NCP=3 (best start offset) is limited to N <= 8192.  As in the
! reference, a loop that ends with an error (INFO < 0) prints its ERROR line and
! still writes FN.YU, FN.ZZ and FN.SY.
PROGRAM jk
  USE, INTRINSIC :: ISO_C_BINDING
  USE, INTRINSIC :: ISO_FORTRAN_ENV, ONLY: OUTPUT_UNIT, ERROR_UNIT
  USE jkg_c
  IMPLICIT NONE
#ifdef ZJK
  CHARACTER(LEN=*), PARAMETER :: USAGE = 'zjk.exe FN N N_2 [ID_NCP [ID_TRU]]'
  COMPLEX(c_double_complex), ALLOCATABLE, TARGET :: Y(:,:), YU(:,:), ZZ(:,:)
#else
  CHARACTER(LEN=*), PARAMETER :: USAGE = 'djk.exe FN N N_2 [ID_NCP [ID_TRU]]'
  REAL(c_double), ALLOCATABLE, TARGET :: Y(:,:), YU(:,:), ZZ(:,:)
#endif
  REAL(c_double), ALLOCATABLE :: SY(:)
  REAL(c_float), ALLOCATABLE, TARGET :: CTL(:)
  INTEGER(c_int64_t), ALLOCATABLE :: J(:)
  CHARACTER(LEN=252) :: FN
  CHARACTER(LEN=1024) :: PATH
  INTEGER :: N, N_2, ID_NCP, ID_TRU, INFO, I, NP, NM, U1, U2, CAP
  INTEGER(c_int) :: RC, DEV
  INTEGER(c_int64_t) :: TT
  TYPE(c_ptr) :: CTX
  TYPE(jkg_opts) :: OPTS
  TYPE(jkg_mixed_opts) :: MOPTS
  TYPE(jkg_result) :: RES
  CHARACTER(LEN=32) :: ENVV, SOLVER

  CALL READCL(FN, N, N_2, ID_NCP, ID_TRU, INFO)
  IF (INFO .NE. 0) ERROR STOP USAGE
  IF (N .LE. 2) ERROR STOP 'N should be at least 3 (for N=2 use dtest)'
  N_2 = jkg_jstep_len(INT(N,c_int), INT(N_2,c_int))
  IF (N_2 .LE. 0) ERROR STOP 'JSTEP_LEN'
  IF ((ID_NCP .LT. 0) .OR. (ID_NCP .GT. 3) .OR. (ID_TRU .LT. 0) .OR. (ID_TRU .GT. 2)) ERROR STOP 'DPROC_INIT'
  SOLVER = 'plain'
  CALL GET_ENVIRONMENT_VARIABLE('JKG_SOLVER', ENVV, STATUS=INFO)
  IF (INFO .EQ. 0) SOLVER = ADJUSTL(ENVV)
  IF ((SOLVER .NE. 'plain') .AND. (SOLVER .NE. 'mixed') .AND. (SOLVER .NE. 'learned') .AND. &
      (SOLVER .NE. 'warm')) ERROR STOP 'JKG_SOLVER: plain, mixed, learned or warm'
  IF ((SOLVER .NE. 'plain') .AND. (ID_NCP .EQ. 3)) ERROR STOP 'JKG_SOLVER: ID_NCP=3 needs the plain solver'

  ALLOCATE(Y(N,N), J(N))
  OPEN(NEWUNIT=U1, FILE=TRIM(FN)//'.Y', STATUS='OLD', ACTION='READ', ACCESS='STREAM', FORM='UNFORMATTED', &
       IOSTAT=INFO)
  IF (INFO .EQ. 0) OPEN(NEWUNIT=U2, FILE=TRIM(FN)//'.J', STATUS='OLD', ACTION='READ', ACCESS='STREAM', &
       FORM='UNFORMATTED', IOSTAT=INFO)
  IF (INFO .NE. 0) THEN
     WRITE (ERROR_UNIT,'(A,I11)') 'INFO=', INFO
     ERROR STOP 'OPEN_YJ_RO'
  END IF
  READ (U1, IOSTAT=INFO) Y
  IF (INFO .EQ. 0) READ (U2, IOSTAT=INFO) J
  IF (INFO .NE. 0) THEN
     WRITE (ERROR_UNIT,'(A,I11)') 'INFO=', INFO
     ERROR STOP 'DREAD_YJ'
  END IF
  CLOSE(U2)
  CLOSE(U1)

  ! the pair census of R%TRU (src/shared/jstep.F90:32-61)
  NP = 0
  NM = 0
  DO I = 1, N
     IF (J(I) .EQ. 1_c_int64_t) THEN
        NP = NP + 1
     ELSE IF (J(I) .EQ. -1_c_int64_t) THEN
        NM = NM + 1
     ELSE
        WRITE (ERROR_UNIT,'(A,I11)') 'INFO=', -2
        ERROR STOP 'R%TRU'
     END IF
  END DO
  TT = (INT(NP,c_int64_t) * (NP - 1)) / 2 + (INT(NM,c_int64_t) * (NM - 1)) / 2
  WRITE (OUTPUT_UNIT,'(3(A,I11),A)') 'J has', NP, ' +  and', NM, ' - signs for <=', TT, ' trig. transfs.'
  FLUSH(OUTPUT_UNIT)

  DEV = 0
  CALL GET_ENVIRONMENT_VARIABLE('JKG_DEVICE', ENVV, STATUS=INFO)
  IF (INFO .EQ. 0) READ (ENVV,*) DEV
  OPTS%fused = 0
  CALL GET_ENVIRONMENT_VARIABLE('JKG_FUSED', ENVV, STATUS=INFO)
  IF (INFO .EQ. 0) READ (ENVV,*) OPTS%fused
  CAP = 64 * N + 64
  ALLOCATE(TRACE(CAP), YU(N,N), ZZ(N,N), SY(N))
  OPTS%id_ncp = ID_NCP
  OPTS%id_tru = ID_TRU
  OPTS%max_steps = -1
  OPTS%steps_per_sync = 0
  OPTS%trace_cap = CAP
  OPTS%trace = C_LOC(TRACE)
  OPTS%progress = C_FUNLOC(on_progress)
  OPTS%user = C_NULL_PTR
  TRACE_UNIT = ERROR_UNIT
  TRACE_PRINTED = 0

  CALL jkg_mixed_opts_default(MOPTS)
  IF (SOLVER .EQ. 'warm') MOPTS%fp32_warm = 1
  IF (SOLVER .EQ. 'learned') THEN
     CALL GET_ENVIRONMENT_VARIABLE('JKG_CTL_PARAMS', PATH, STATUS=INFO)
     IF (INFO .NE. 0) THEN
        CALL GET_COMMAND_ARGUMENT(0, PATH)
        I = INDEX(PATH, '/', BACK=.TRUE.)
        PATH = PATH(1:I)//'../jkogb_amd/ctl_params.f32'
     END IF
     ALLOCATE(CTL(JKG_CTL_NPARAM))
     OPEN(NEWUNIT=U1, FILE=TRIM(PATH), STATUS='OLD', ACTION='READ', ACCESS='STREAM', FORM='UNFORMATTED', &
          IOSTAT=INFO)
     IF (INFO .EQ. 0) READ (U1, IOSTAT=INFO) CTL
     IF (INFO .NE. 0) THEN
        WRITE (ERROR_UNIT,'(A,I11,2A)') 'INFO=', INFO, ' ', TRIM(PATH)
        ERROR STOP 'JKG_CTL_PARAMS'
     END IF
     CLOSE(U1)
     MOPTS%controller = 1
     MOPTS%ctl = C_LOC(CTL)
  END IF

  RC = jkg_create(CTX, DEV)
  IF (RC .NE. JKG_OK) THEN
     WRITE (ERROR_UNIT,'(A,I11,2A)') 'INFO=', RC, ' ', TRIM(jkg_message(RC))
     ERROR STOP 'jkg_create'
  END IF
  CALL write_trace_header(ERROR_UNIT)
#ifdef ZJK
  IF (SOLVER .EQ. 'plain') THEN
     RC = jkg_zsolve(CTX, INT(N,c_int), INT(N_2,c_int), Y, J, YU, ZZ, SY, OPTS, RES)
  ELSE
     RC = jkg_zsolve_mixed(CTX, INT(N,c_int), INT(N_2,c_int), Y, J, YU, ZZ, SY, OPTS, MOPTS, RES)
  END IF
#else
  IF (SOLVER .EQ. 'plain') THEN
     RC = jkg_dsolve(CTX, INT(N,c_int), INT(N_2,c_int), Y, J, YU, ZZ, SY, OPTS, RES)
  ELSE
     RC = jkg_dsolve_mixed(CTX, INT(N,c_int), INT(N_2,c_int), Y, J, YU, ZZ, SY, OPTS, MOPTS, RES)
  END IF
#endif
  CALL jkg_destroy(CTX)
  IF (RC .NE. JKG_OK) THEN
     WRITE (ERROR_UNIT,'(A,I11,2A)') 'INFO=', RC, ' ', TRIM(jkg_message(RC))
     ERROR STOP 'DSTEP_LOOP'
  END IF
  CALL write_trace_upto(MIN(INT(RES%steps), CAP))
  IF (RES%info .GE. 0) THEN
     WRITE (OUTPUT_UNIT,'(A,I11,A,F13.6,A)') 'Executed ', RES%info, ' steps with transformations in ', &
          RES%seconds, ' s'
     FLUSH(OUTPUT_UNIT)
  ELSE ! the loop's error: reported, and the factors written all the same (src/D/djk.F90:86-114)
     WRITE (ERROR_UNIT,'(A,I11,A,F13.6,A)') 'ERROR ', RES%info, ' after ', RES%seconds, ' s'
     FLUSH(ERROR_UNIT)
  END IF

  OPEN(NEWUNIT=U1, FILE=TRIM(FN)//'.YU', STATUS='REPLACE', ACTION='WRITE', ACCESS='STREAM', &
       FORM='UNFORMATTED', IOSTAT=INFO)
  IF (INFO .EQ. 0) WRITE (U1, IOSTAT=INFO) YU
  IF (INFO .EQ. 0) CLOSE(U1)
  IF (INFO .EQ. 0) OPEN(NEWUNIT=U1, FILE=TRIM(FN)//'.ZZ', STATUS='REPLACE', ACTION='WRITE', ACCESS='STREAM', &
       FORM='UNFORMATTED', IOSTAT=INFO)
  IF (INFO .EQ. 0) WRITE (U1, IOSTAT=INFO) ZZ
  IF (INFO .EQ. 0) CLOSE(U1)
  IF (INFO .EQ. 0) OPEN(NEWUNIT=U1, FILE=TRIM(FN)//'.SY', STATUS='REPLACE', ACTION='WRITE', ACCESS='STREAM', &
       FORM='UNFORMATTED', IOSTAT=INFO)
  IF (INFO .EQ. 0) WRITE (U1, IOSTAT=INFO) SY
  IF (INFO .EQ. 0) CLOSE(U1)
  IF (INFO .NE. 0) THEN
     WRITE (ERROR_UNIT,'(A,I11)') 'INFO=', INFO
     ERROR STOP 'DWRITE_UZS'
  END IF

CONTAINS

  ! positional command line, READCL semantics (src/shared/readcl.F90:1-85)
  SUBROUTINE READCL(FN, N, N_2, ID_NCP, ID_TRU, INFO)
    CHARACTER(LEN=*), INTENT(OUT) :: FN
    INTEGER, INTENT(OUT) :: N, N_2, ID_NCP, ID_TRU, INFO
    CHARACTER(LEN=252) :: ARG
    INTEGER :: ARGC, IOS
    N = 0
    N_2 = 0
    ID_NCP = 0
    ID_TRU = 0
    ARGC = COMMAND_ARGUMENT_COUNT()
    IF (ARGC .LT. 2) THEN
       INFO = -6
       RETURN
    ELSE IF (ARGC .GT. 5) THEN
       INFO = 6
       RETURN
    END IF
    INFO = 0
    CALL GET_COMMAND_ARGUMENT(1, FN)
    IF (LEN_TRIM(FN) .LE. 0) THEN
       INFO = 1
       RETURN
    END IF
    CALL GET_COMMAND_ARGUMENT(2, ARG)
    READ (ARG,*,IOSTAT=IOS) N
    IF ((IOS .NE. 0) .OR. (N .LT. 0)) THEN
       INFO = 2
       RETURN
    END IF
    IF (ARGC .GE. 3) THEN
       CALL GET_COMMAND_ARGUMENT(3, ARG)
       READ (ARG,*,IOSTAT=IOS) N_2
       IF ((IOS .NE. 0) .OR. (N_2 .LT. 0)) THEN
          INFO = 3
          RETURN
       END IF
    END IF
    IF (ARGC .GE. 4) THEN
       CALL GET_COMMAND_ARGUMENT(4, ARG)
       READ (ARG,*,IOSTAT=IOS) ID_NCP
       IF ((IOS .NE. 0) .OR. (ID_NCP .LT. 0)) THEN
          INFO = 4
          RETURN
       END IF
    END IF
    IF (ARGC .GE. 5) THEN
       CALL GET_COMMAND_ARGUMENT(5, ARG)
       READ (ARG,*,IOSTAT=IOS) ID_TRU
       IF ((IOS .NE. 0) .OR. (ID_TRU .LT. 0)) THEN
          INFO = 5
          RETURN
       END IF
    END IF
  END SUBROUTINE READCL
END PROGRAM jk
