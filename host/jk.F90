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
!   warm_learned  both: FP32 steps first, the trained controller throughout
!           (real problems only)
! Other environment: JKG_DEVICE (default 0), JKG_FUSED=1 selects FMAD = fma.
!
! Checkpoint and resume (plain solver; jkg_dsolve_from / jkg_zsolve_from):
!   JKG_CHECKPOINT=K  the loop stops every K steps and writes its state: FN.CA,
!                     FN.CU, FN.CZ (A, U and Z as the loop holds them, N x N
!                     column-major, the element type of FN.Y) and FN.CS (the
!                     next step's index, one int64; -1 while the others are
!                     being written), then goes on; the outputs, stdout and the
!                     CSV are those of one uninterrupted run
!   JKG_RESUME=1      starts from FN.CA, FN.CU, FN.CZ at step FN.CS instead of
!                     FN.Y (FN.J is still read); the CSV continues from that
!                     step and the last stdout line counts all steps (its time:
!                     this run's)
! Frames (the role of the reference's ANIMATE build, src/D/dstep.F90:485-592):
!   JKG_FRAMES=E[,S[,M]]  a S x S byte picture of |A| at the start of every step
!                     s = 0 mod E (jkg_set_frames; S = 64, at most M = 1024 per
!                     solve call), appended row-major to FN.FRM
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
  COMPLEX(c_double_complex), ALLOCATABLE, TARGET :: Y(:,:), YU(:,:), ZZ(:,:), CU(:,:), CZ(:,:)
  COMPLEX(c_double_complex), PARAMETER :: ONE = (1.0_c_double, 0.0_c_double)
#else
  CHARACTER(LEN=*), PARAMETER :: USAGE = 'djk.exe FN N N_2 [ID_NCP [ID_TRU]]'
  REAL(c_double), ALLOCATABLE, TARGET :: Y(:,:), YU(:,:), ZZ(:,:), CU(:,:), CZ(:,:)
  REAL(c_double), PARAMETER :: ONE = 1.0_c_double
#endif
  REAL(c_double), ALLOCATABLE :: SY(:)
  REAL(c_float), ALLOCATABLE, TARGET :: CTL(:)
  INTEGER(c_int64_t), ALLOCATABLE :: J(:)
  INTEGER(c_int8_t), ALLOCATABLE :: FRM(:)
  INTEGER(c_int32_t) :: NFRM
  CHARACTER(LEN=252) :: FN
  CHARACTER(LEN=1024) :: PATH
  INTEGER :: N, N_2, ID_NCP, ID_TRU, INFO, I, NP, NM, U1, U2, CAP
  INTEGER :: CKPT, RESUME, FEVERY, FSIZE, FMAX, UF, STEP0
  INTEGER(c_int) :: RC, DEV
  INTEGER(c_int64_t) :: TT, S64
  REAL(c_double) :: SECS
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
      (SOLVER .NE. 'warm') .AND. (SOLVER .NE. 'warm_learned')) &
       ERROR STOP 'JKG_SOLVER: plain, mixed, learned, warm or warm_learned'
  IF ((SOLVER .NE. 'plain') .AND. (ID_NCP .EQ. 3)) ERROR STOP 'JKG_SOLVER: ID_NCP=3 needs the plain solver'
  CKPT = 0
  CALL GET_ENVIRONMENT_VARIABLE('JKG_CHECKPOINT', ENVV, STATUS=INFO)
  IF (INFO .EQ. 0) THEN
     READ (ENVV,*,IOSTAT=INFO) CKPT
     IF ((INFO .NE. 0) .OR. (CKPT .LE. 0)) ERROR STOP 'JKG_CHECKPOINT: a positive number of steps'
  END IF
  RESUME = 0
  CALL GET_ENVIRONMENT_VARIABLE('JKG_RESUME', ENVV, STATUS=INFO)
  IF (INFO .EQ. 0) THEN
     READ (ENVV,*,IOSTAT=INFO) RESUME
     IF ((INFO .NE. 0) .OR. (RESUME .LT. 0) .OR. (RESUME .GT. 1)) ERROR STOP 'JKG_RESUME: 0 or 1'
  END IF
  IF (((CKPT .GT. 0) .OR. (RESUME .EQ. 1)) .AND. (SOLVER .NE. 'plain')) &
       ERROR STOP 'JKG_CHECKPOINT, JKG_RESUME: the plain solver only'
  FEVERY = 0
  FSIZE = 64
  FMAX = 1024
  CALL GET_ENVIRONMENT_VARIABLE('JKG_FRAMES', ENVV, STATUS=INFO)
  IF (INFO .EQ. 0) THEN
     PATH = TRIM(ENVV)//' /'  ! '/' ends the list: absent values keep their defaults
     READ (PATH,*,IOSTAT=INFO) FEVERY, FSIZE, FMAX
     IF ((INFO .NE. 0) .OR. (FEVERY .LE. 0) .OR. (FSIZE .LE. 0) .OR. (FMAX .LE. 0)) &
          ERROR STOP 'JKG_FRAMES: E[,S[,M]], positive'
  END IF

  ALLOCATE(Y(N,N), J(N))
  STEP0 = 0
  IF (RESUME .EQ. 1) THEN  ! the checkpoint in place of FN.Y
     ALLOCATE(CU(N,N), CZ(N,N))
     CALL read_state(TRIM(FN)//'.CA', Y)
     CALL read_state(TRIM(FN)//'.CU', CU)
     CALL read_state(TRIM(FN)//'.CZ', CZ)
     OPEN(NEWUNIT=U1, FILE=TRIM(FN)//'.CS', STATUS='OLD', ACTION='READ', ACCESS='STREAM', FORM='UNFORMATTED', &
          IOSTAT=INFO)
     IF (INFO .EQ. 0) READ (U1, IOSTAT=INFO) S64
     IF ((INFO .NE. 0) .OR. (S64 .LT. 0_c_int64_t) .OR. (S64 .GT. INT(HUGE(STEP0),c_int64_t))) THEN
        WRITE (ERROR_UNIT,'(A,I11)') 'INFO=', INFO
        ERROR STOP 'READ_CHECKPOINT'
     END IF
     CLOSE(U1)
     STEP0 = INT(S64)
     OPEN(NEWUNIT=U2, FILE=TRIM(FN)//'.J', STATUS='OLD', ACTION='READ', ACCESS='STREAM', FORM='UNFORMATTED', &
          IOSTAT=INFO)
     IF (INFO .NE. 0) THEN
        WRITE (ERROR_UNIT,'(A,I11)') 'INFO=', INFO
        ERROR STOP 'OPEN_YJ_RO'
     END IF
     READ (U2, IOSTAT=INFO) J
     IF (INFO .NE. 0) THEN
        WRITE (ERROR_UNIT,'(A,I11)') 'INFO=', INFO
        ERROR STOP 'DREAD_YJ'
     END IF
     CLOSE(U2)
  ELSE
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
     IF (CKPT .GT. 0) THEN  ! the loop's own start: U = I, Z = I
        ALLOCATE(CU(N,N), CZ(N,N))
        CU = 0
        CZ = 0
        DO I = 1, N
           CU(I,I) = ONE
           CZ(I,I) = ONE
        END DO
     END IF
  END IF

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
  IF ((SOLVER .EQ. 'warm') .OR. (SOLVER .EQ. 'warm_learned')) MOPTS%fp32_warm = 1
  IF ((SOLVER .EQ. 'learned') .OR. (SOLVER .EQ. 'warm_learned')) THEN
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
  IF (FEVERY .GT. 0) THEN
     ALLOCATE(FRM(INT(FSIZE,c_int64_t) * FSIZE * FMAX))
     RC = jkg_set_frames(CTX, INT(FEVERY,c_int), INT(FSIZE,c_int), INT(FMAX,c_int), FRM, NFRM)
     IF (RC .NE. JKG_OK) THEN
        WRITE (ERROR_UNIT,'(A,I11,2A)') 'INFO=', RC, ' ', TRIM(jkg_message(RC))
        ERROR STOP 'JKG_FRAMES'
     END IF
     OPEN(NEWUNIT=UF, FILE=TRIM(FN)//'.FRM', STATUS='REPLACE', ACTION='WRITE', ACCESS='STREAM', &
          FORM='UNFORMATTED', IOSTAT=INFO)
     IF (INFO .NE. 0) ERROR STOP 'OPEN_FRM'
  END IF
  CALL write_trace_header(ERROR_UNIT)
  IF (ALLOCATED(CU)) THEN  ! the loop in parts, from and to its state
     SECS = 0.0_c_double
     DO
        IF (CKPT .GT. 0) OPTS%max_steps = CKPT
        TRACE_PRINTED = 0
#ifdef ZJK
        RC = jkg_zsolve_from(CTX, INT(N,c_int), INT(N_2,c_int), Y, CU, CZ, J, INT(STEP0,c_int), YU, ZZ, SY, &
             OPTS, RES)
#else
        RC = jkg_dsolve_from(CTX, INT(N,c_int), INT(N_2,c_int), Y, CU, CZ, J, INT(STEP0,c_int), YU, ZZ, SY, &
             OPTS, RES)
#endif
        IF (RC .NE. JKG_OK) EXIT
        CALL write_trace_upto(MIN(INT(RES%steps), CAP))
        CALL write_frames()
        SECS = SECS + RES%seconds
        ! ended: no pivot, no effective transform, or a failed 2x2 (INFO < 0)
        IF ((CKPT .LE. 0) .OR. (RES%info .LT. STEP0 + CKPT)) EXIT
        STEP0 = RES%info
        CALL write_checkpoint()
     END DO
     RES%seconds = SECS
  ELSE
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
     IF (RC .EQ. JKG_OK) THEN
        CALL write_trace_upto(MIN(INT(RES%steps), CAP))
        CALL write_frames()
     END IF
  END IF
  CALL jkg_destroy(CTX)
  IF (RC .NE. JKG_OK) THEN
     WRITE (ERROR_UNIT,'(A,I11,2A)') 'INFO=', RC, ' ', TRIM(jkg_message(RC))
     ERROR STOP 'DSTEP_LOOP'
  END IF
  IF (FEVERY .GT. 0) CLOSE(UF)
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

  ! an N x N state matrix of a checkpoint
  SUBROUTINE read_state(NAME, X)
    CHARACTER(LEN=*), INTENT(IN) :: NAME
#ifdef ZJK
    COMPLEX(c_double_complex), INTENT(OUT) :: X(:,:)
#else
    REAL(c_double), INTENT(OUT) :: X(:,:)
#endif
    INTEGER :: U, IOS
    OPEN(NEWUNIT=U, FILE=NAME, STATUS='OLD', ACTION='READ', ACCESS='STREAM', FORM='UNFORMATTED', IOSTAT=IOS)
    IF (IOS .EQ. 0) READ (U, IOSTAT=IOS) X
    IF (IOS .NE. 0) THEN
       WRITE (ERROR_UNIT,'(A,I11,2A)') 'INFO=', IOS, ' ', NAME
       ERROR STOP 'READ_CHECKPOINT'
    END IF
    CLOSE(U)
  END SUBROUTINE read_state

  ! FN.CS := -1, FN.CA, FN.CU, FN.CZ := the loop's state, FN.CS := STEP0
  SUBROUTINE write_checkpoint()
    INTEGER :: U, IOS
    CALL write_step(-1_c_int64_t)
    OPEN(NEWUNIT=U, FILE=TRIM(FN)//'.CA', STATUS='REPLACE', ACTION='WRITE', ACCESS='STREAM', &
         FORM='UNFORMATTED', IOSTAT=IOS)
    IF (IOS .EQ. 0) WRITE (U, IOSTAT=IOS) Y
    IF (IOS .EQ. 0) CLOSE(U)
    IF (IOS .EQ. 0) OPEN(NEWUNIT=U, FILE=TRIM(FN)//'.CU', STATUS='REPLACE', ACTION='WRITE', ACCESS='STREAM', &
         FORM='UNFORMATTED', IOSTAT=IOS)
    IF (IOS .EQ. 0) WRITE (U, IOSTAT=IOS) CU
    IF (IOS .EQ. 0) CLOSE(U)
    IF (IOS .EQ. 0) OPEN(NEWUNIT=U, FILE=TRIM(FN)//'.CZ', STATUS='REPLACE', ACTION='WRITE', ACCESS='STREAM', &
         FORM='UNFORMATTED', IOSTAT=IOS)
    IF (IOS .EQ. 0) WRITE (U, IOSTAT=IOS) CZ
    IF (IOS .EQ. 0) CLOSE(U)
    IF (IOS .NE. 0) THEN
       WRITE (ERROR_UNIT,'(A,I11)') 'INFO=', IOS
       ERROR STOP 'WRITE_CHECKPOINT'
    END IF
    CALL write_step(INT(STEP0,c_int64_t))
  END SUBROUTINE write_checkpoint

  SUBROUTINE write_step(S)
    INTEGER(c_int64_t), INTENT(IN) :: S
    INTEGER :: U, IOS
    OPEN(NEWUNIT=U, FILE=TRIM(FN)//'.CS', STATUS='REPLACE', ACTION='WRITE', ACCESS='STREAM', &
         FORM='UNFORMATTED', IOSTAT=IOS)
    IF (IOS .EQ. 0) WRITE (U, IOSTAT=IOS) S
    IF (IOS .EQ. 0) CLOSE(U)
    IF (IOS .NE. 0) THEN
       WRITE (ERROR_UNIT,'(A,I11)') 'INFO=', IOS
       ERROR STOP 'WRITE_CHECKPOINT'
    END IF
  END SUBROUTINE write_step

  ! the frames of the last solve call, appended to FN.FRM
  SUBROUTINE write_frames()
    INTEGER :: IOS
    IF (FEVERY .LE. 0) RETURN
    IF (NFRM .LE. 0) RETURN
    WRITE (UF, IOSTAT=IOS) FRM(1:INT(NFRM,c_int64_t) * FSIZE * FSIZE)
    IF (IOS .NE. 0) ERROR STOP 'WRITE_FRM'
  END SUBROUTINE write_frames

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
