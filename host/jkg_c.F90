! jkg_c.F90 -- ISO_C_BINDING interface to libjkg.so (include/jkg.h) for the
! Fortran drop-in programs djk.exe and zjk.exe.  The derived types mirror the
! C structs field by field (BIND(C) gives them the C layout).
MODULE jkg_c
  USE, INTRINSIC :: ISO_C_BINDING
  IMPLICIT NONE

  INTEGER(c_int), PARAMETER :: JKG_OK = 0_c_int
  INTEGER(c_int), PARAMETER :: JKG_CTL_NPARAM = 5378_c_int

  TYPE, BIND(C) :: jkg_steprec
     INTEGER(c_int32_t) :: step, oldlen, newlen, trig, hyp, it
     REAL(c_double) :: tw
     REAL(c_float) :: build_s, transf_s
  END TYPE jkg_steprec

  TYPE, BIND(C) :: jkg_opts
     INTEGER(c_int32_t) :: id_ncp, id_tru, fused, max_steps, steps_per_sync, trace_cap
     TYPE(c_ptr) :: trace
     TYPE(c_funptr) :: progress = c_null_funptr
     TYPE(c_ptr) :: user = c_null_ptr
  END TYPE jkg_opts

  TYPE, BIND(C) :: jkg_result
     INTEGER(c_int32_t) :: info, tt, npos, nneg, n2, steps
     REAL(c_double) :: seconds
     INTEGER(c_int64_t) :: sum_sl, sum_it, sum_b2, sum_b4, sum_abod
  END TYPE jkg_result

  ! jkg_mixed_opts: the options of the mixed solver (see include/jkg.h)
  TYPE, BIND(C) :: jkg_mixed_opts
     INTEGER(c_int32_t) :: weights, fp32_warm
     REAL(c_double) :: switch_tol
     INTEGER(c_int32_t) :: key_bits, controller
     TYPE(c_ptr) :: ctl
     REAL(c_double) :: threshold, tol
     TYPE(c_ptr) :: ctl_trace
     INTEGER(c_int32_t) :: ctl_trace_cap, ordering
     TYPE(c_ptr) :: warm
     REAL(c_double) :: ctl_tau_max
     REAL(c_float) :: ctl_handoff
     INTEGER(c_int32_t) :: reserved
  END TYPE jkg_mixed_opts

  INTERFACE
     INTEGER(c_int) FUNCTION jkg_create(ctx, device) BIND(C, NAME='jkg_create')
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), INTENT(OUT) :: ctx
       INTEGER(c_int), VALUE :: device
     END FUNCTION jkg_create

     SUBROUTINE jkg_destroy(ctx) BIND(C, NAME='jkg_destroy')
       IMPORT :: c_ptr
       TYPE(c_ptr), VALUE :: ctx
     END SUBROUTINE jkg_destroy

     TYPE(c_ptr) FUNCTION jkg_strerror(code) BIND(C, NAME='jkg_strerror')
       IMPORT :: c_int, c_ptr
       INTEGER(c_int), VALUE :: code
     END FUNCTION jkg_strerror

     INTEGER(c_int) FUNCTION jkg_jstep_len(n, n2) BIND(C, NAME='jkg_jstep_len')
       IMPORT :: c_int
       INTEGER(c_int), VALUE :: n, n2
     END FUNCTION jkg_jstep_len

     SUBROUTINE jkg_mixed_opts_default(m) BIND(C, NAME='jkg_mixed_opts_default')
       IMPORT :: jkg_mixed_opts
       TYPE(jkg_mixed_opts), INTENT(OUT) :: m
     END SUBROUTINE jkg_mixed_opts_default

     INTEGER(c_int) FUNCTION jkg_dsolve(ctx, n, n2, y, j, yu, zz, sy, opts, res) BIND(C, NAME='jkg_dsolve')
       IMPORT :: c_int, c_ptr, c_double, c_int64_t, jkg_opts, jkg_result
       TYPE(c_ptr), VALUE :: ctx
       INTEGER(c_int), VALUE :: n, n2
       REAL(c_double), INTENT(IN) :: y(*)
       INTEGER(c_int64_t), INTENT(IN) :: j(*)
       REAL(c_double), INTENT(OUT) :: yu(*), zz(*), sy(*)
       TYPE(jkg_opts), INTENT(IN) :: opts
       TYPE(jkg_result), INTENT(OUT) :: res
     END FUNCTION jkg_dsolve

     ! the loop resumed from its state (a, u, z) at step step0; (a, u, z) returned after the last step
     INTEGER(c_int) FUNCTION jkg_dsolve_from(ctx, n, n2, a, u, z, j, step0, yu, zz, sy, opts, res) &
          BIND(C, NAME='jkg_dsolve_from')
       IMPORT :: c_int, c_ptr, c_double, c_int64_t, jkg_opts, jkg_result
       TYPE(c_ptr), VALUE :: ctx
       INTEGER(c_int), VALUE :: n, n2, step0
       REAL(c_double), INTENT(INOUT) :: a(*), u(*), z(*)
       INTEGER(c_int64_t), INTENT(IN) :: j(*)
       REAL(c_double), INTENT(OUT) :: yu(*), zz(*), sy(*)
       TYPE(jkg_opts), INTENT(IN) :: opts
       TYPE(jkg_result), INTENT(OUT) :: res
     END FUNCTION jkg_dsolve_from

     INTEGER(c_int) FUNCTION jkg_zsolve_from(ctx, n, n2, a, u, z, j, step0, yu, zz, sy, opts, res) &
          BIND(C, NAME='jkg_zsolve_from')
       IMPORT :: c_int, c_ptr, c_double, c_double_complex, c_int64_t, jkg_opts, jkg_result
       TYPE(c_ptr), VALUE :: ctx
       INTEGER(c_int), VALUE :: n, n2, step0
       COMPLEX(c_double_complex), INTENT(INOUT) :: a(*), u(*), z(*)
       INTEGER(c_int64_t), INTENT(IN) :: j(*)
       COMPLEX(c_double_complex), INTENT(OUT) :: yu(*), zz(*)
       REAL(c_double), INTENT(OUT) :: sy(*)
       TYPE(jkg_opts), INTENT(IN) :: opts
       TYPE(jkg_result), INTENT(OUT) :: res
     END FUNCTION jkg_zsolve_from

     ! pictures of |A| during later solves (size x size bytes every `every` steps)
     INTEGER(c_int) FUNCTION jkg_set_frames(ctx, every, fsize, max_frames, frames, count) &
          BIND(C, NAME='jkg_set_frames')
       IMPORT :: c_int, c_ptr, c_int8_t, c_int32_t
       TYPE(c_ptr), VALUE :: ctx
       INTEGER(c_int), VALUE :: every, fsize, max_frames
       INTEGER(c_int8_t), INTENT(OUT) :: frames(*)
       INTEGER(c_int32_t), INTENT(OUT) :: count
     END FUNCTION jkg_set_frames

     INTEGER(c_int) FUNCTION jkg_dsolve_mixed(ctx, n, n2, y, j, yu, zz, sy, opts, m, res) &
          BIND(C, NAME='jkg_dsolve_mixed')
       IMPORT :: c_int, c_ptr, c_double, c_int64_t, jkg_opts, jkg_mixed_opts, jkg_result
       TYPE(c_ptr), VALUE :: ctx
       INTEGER(c_int), VALUE :: n, n2
       REAL(c_double), INTENT(IN) :: y(*)
       INTEGER(c_int64_t), INTENT(IN) :: j(*)
       REAL(c_double), INTENT(OUT) :: yu(*), zz(*), sy(*)
       TYPE(jkg_opts), INTENT(IN) :: opts
       TYPE(jkg_mixed_opts), INTENT(IN) :: m
       TYPE(jkg_result), INTENT(OUT) :: res
     END FUNCTION jkg_dsolve_mixed

     INTEGER(c_int) FUNCTION jkg_zsolve(ctx, n, n2, y, j, yu, zz, sy, opts, res) BIND(C, NAME='jkg_zsolve')
       IMPORT :: c_int, c_ptr, c_double, c_double_complex, c_int64_t, jkg_opts, jkg_result
       TYPE(c_ptr), VALUE :: ctx
       INTEGER(c_int), VALUE :: n, n2
       COMPLEX(c_double_complex), INTENT(IN) :: y(*)
       INTEGER(c_int64_t), INTENT(IN) :: j(*)
       COMPLEX(c_double_complex), INTENT(OUT) :: yu(*), zz(*)
       REAL(c_double), INTENT(OUT) :: sy(*)
       TYPE(jkg_opts), INTENT(IN) :: opts
       TYPE(jkg_result), INTENT(OUT) :: res
     END FUNCTION jkg_zsolve

     INTEGER(c_int) FUNCTION jkg_zsolve_mixed(ctx, n, n2, y, j, yu, zz, sy, opts, m, res) &
          BIND(C, NAME='jkg_zsolve_mixed')
       IMPORT :: c_int, c_ptr, c_double, c_double_complex, c_int64_t, jkg_opts, jkg_mixed_opts, jkg_result
       TYPE(c_ptr), VALUE :: ctx
       INTEGER(c_int), VALUE :: n, n2
       COMPLEX(c_double_complex), INTENT(IN) :: y(*)
       INTEGER(c_int64_t), INTENT(IN) :: j(*)
       COMPLEX(c_double_complex), INTENT(OUT) :: yu(*), zz(*)
       REAL(c_double), INTENT(OUT) :: sy(*)
       TYPE(jkg_opts), INTENT(IN) :: opts
       TYPE(jkg_mixed_opts), INTENT(IN) :: m
       TYPE(jkg_result), INTENT(OUT) :: res
     END FUNCTION jkg_zsolve_mixed
  END INTERFACE

  ! The per-step CSV of the reference (src/D/dstep.F90:388-440, 559-562), printed
  ! as the steps come: the trace buffer the solve fills and the lines printed.
  TYPE(jkg_steprec), ALLOCATABLE, TARGET :: TRACE(:)
  INTEGER :: TRACE_PRINTED = 0
  INTEGER :: TRACE_UNIT = 0

CONTAINS

  ! the message of a status code, as a Fortran string
  FUNCTION jkg_message(code) RESULT(msg)
    INTEGER(c_int), INTENT(IN) :: code
    CHARACTER(LEN=80) :: msg
    CHARACTER(KIND=c_char), POINTER :: s(:)
    INTEGER :: i
    CALL C_F_POINTER(jkg_strerror(code), s, [80])
    msg = ' '
    DO i = 1, 80
       IF (s(i) .EQ. C_NULL_CHAR) EXIT
       msg(i:i) = s(i)
    END DO
  END FUNCTION jkg_message

  ! the CSV header line of the reference (src/D/dstep.F90:559-562)
  SUBROUTINE write_trace_header(u)
    INTEGER, INTENT(IN) :: u
    WRITE (u,'(A)') '"STEP","OLDLEN","BUILDs","TRANSFs","NEWLEN","NORMF2D","TRIG","HYP"'
    FLUSH(u)
  END SUBROUTINE write_trace_header

  ! CSV lines of the records TRACE(TRACE_PRINTED+1 : upto), in the reference's
  ! formats (src/D/dstep.F90:388-440); BUILDs and TRANSFs are the device times
  SUBROUTINE write_trace_upto(upto)
    INTEGER, INTENT(IN) :: upto
    INTEGER :: k
    DO k = TRACE_PRINTED + 1, MIN(upto, SIZE(TRACE))
       WRITE (TRACE_UNIT,'(I11,A,I11,A,F12.6,A,F12.6,A,I11,A,ES25.17E3,2(A,I11))') TRACE(k)%step, ',', &
            TRACE(k)%oldlen, ',', REAL(TRACE(k)%build_s, c_double), ',', REAL(TRACE(k)%transf_s, c_double), ',', &
            TRACE(k)%newlen, ',', TRACE(k)%tw, ',', TRACE(k)%trig, ',', TRACE(k)%hyp
       TRACE_PRINTED = k
    END DO
    FLUSH(TRACE_UNIT)
  END SUBROUTINE write_trace_upto

  ! progress hook of the solve: the library has copied the records of the
  ! steps done so far into TRACE before calling it
  SUBROUTINE on_progress(user, steps, done) BIND(C)
    TYPE(c_ptr), VALUE :: user
    INTEGER(c_int32_t), VALUE :: steps, done
    CALL write_trace_upto(INT(steps))
  END SUBROUTINE on_progress
END MODULE jkg_c
