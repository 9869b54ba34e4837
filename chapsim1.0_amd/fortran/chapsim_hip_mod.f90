!-----------------------------------------------------------------------------------------------
! chapsim_hip: ISO_C_BINDING interface to libchapsim_hip.so (include/chapsim_hip.h), the
! MI355X implementation of the RK3/Crank-Nicolson substep SOLVERRK3_MOM_tg
! (SOLVERRK3_MOM_tg.f90:13-56).  Every function returns 0 on success or a negative CHS_E_*
! code; chs_error_message() returns the library's message for a context.
!
! Arrays are passed as C_LOC of the reference's own arrays, which already have the layout
! the library expects (MEMALLOCT.f90:96-103): Q_tg(NCL1,0:NCL2+1,NCL3,3),
! PR_tg(NCL1,0:NCL2+1,NCL3), CONVH0_tg(NCL1,NCL2,NCL3,3), all REAL(8).
!-----------------------------------------------------------------------------------------------
MODULE chapsim_hip
    USE, INTRINSIC :: ISO_C_BINDING
    IMPLICIT NONE

    INTEGER(C_INT), PARAMETER :: CHS_OK = 0
    INTEGER(C_INT), PARAMETER :: CHS_PLAIN_FP64 = 0
    INTEGER(C_INT), PARAMETER :: CHS_MIXED = 1

    ! mirror of chs_coeffs
    TYPE, BIND(C) :: chs_coeffs
        INTEGER(C_INT) :: L, M, N
        REAL(C_DOUBLE) :: dxi, dzi, dxqi, dzqi, volm
        TYPE(C_PTR) :: dyfi, dyci, wff, wfb
        TYPE(C_PTR) :: amvr, acvr, apvr
        TYPE(C_PTR) :: amph, acph, apph
        TYPE(C_PTR) :: ak1, ak3
        TYPE(C_PTR) :: tgam, troh, talp
    END TYPE chs_coeffs

    INTERFACE
        FUNCTION chs_version() BIND(C, NAME='chs_version') RESULT(p)
            IMPORT :: C_PTR
            TYPE(C_PTR) :: p
        END FUNCTION

        FUNCTION chs_last_error(ctx) BIND(C, NAME='chs_last_error') RESULT(p)
            IMPORT :: C_PTR
            TYPE(C_PTR), VALUE :: ctx
            TYPE(C_PTR) :: p
        END FUNCTION

        FUNCTION chs_channel_setup(L, M, N, HX, HZ, STR2, scal, dyfi, dyci, wff, wfb, amvr, acvr, &
                                   apvr, amph, acph, apph, ak1, ak3, tgam, troh, talp) &
                 BIND(C, NAME='chs_channel_setup') RESULT(rc)
            IMPORT :: C_INT, C_DOUBLE
            INTEGER(C_INT), VALUE :: L, M, N
            REAL(C_DOUBLE), VALUE :: HX, HZ, STR2
            REAL(C_DOUBLE) :: scal(5), dyfi(M), dyci(M + 1), wff(M + 1), wfb(M + 1)
            REAL(C_DOUBLE) :: amvr(M + 1, 3), acvr(M + 1, 3), apvr(M + 1, 3)
            REAL(C_DOUBLE) :: amph(M), acph(M), apph(M), ak1(L), ak3(N)
            REAL(C_DOUBLE) :: tgam(0:3), troh(0:3), talp(0:3)
            INTEGER(C_INT) :: rc
        END FUNCTION

        FUNCTION chs_create(c, mode, ctx) BIND(C, NAME='chs_create') RESULT(rc)
            IMPORT :: C_INT, C_PTR, chs_coeffs
            TYPE(chs_coeffs), INTENT(IN) :: c
            INTEGER(C_INT), VALUE :: mode
            TYPE(C_PTR), INTENT(OUT) :: ctx
            INTEGER(C_INT) :: rc
        END FUNCTION

        SUBROUTINE chs_destroy(ctx) BIND(C, NAME='chs_destroy')
            IMPORT :: C_PTR
            TYPE(C_PTR), VALUE :: ctx
        END SUBROUTINE

        FUNCTION chs_upload_fields(ctx, q, pr, convh0) BIND(C, NAME='chs_upload_fields') RESULT(rc)
            IMPORT :: C_INT, C_PTR
            TYPE(C_PTR), VALUE :: ctx, q, pr, convh0
            INTEGER(C_INT) :: rc
        END FUNCTION

        FUNCTION chs_download_fields(ctx, q, pr, convh0) BIND(C, NAME='chs_download_fields') RESULT(rc)
            IMPORT :: C_INT, C_PTR
            TYPE(C_PTR), VALUE :: ctx, q, pr, convh0
            INTEGER(C_INT) :: rc
        END FUNCTION

        FUNCTION chs_substep_tg(ctx, ns, dt, cvisc) BIND(C, NAME='chs_substep_tg') RESULT(rc)
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: ctx
            INTEGER(C_INT), VALUE :: ns
            REAL(C_DOUBLE), VALUE :: dt, cvisc
            INTEGER(C_INT) :: rc
        END FUNCTION

        FUNCTION chs_step_tg(ctx, dt, cvisc) BIND(C, NAME='chs_step_tg') RESULT(rc)
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: ctx
            REAL(C_DOUBLE), VALUE :: dt, cvisc
            INTEGER(C_INT) :: rc
        END FUNCTION

        FUNCTION chs_set_graph(ctx, enable) BIND(C, NAME='chs_set_graph') RESULT(rc)
            IMPORT :: C_INT, C_PTR
            TYPE(C_PTR), VALUE :: ctx
            INTEGER(C_INT), VALUE :: enable
            INTEGER(C_INT) :: rc
        END FUNCTION

        FUNCTION chs_poisson(ctx, rhs, phi) BIND(C, NAME='chs_poisson') RESULT(rc)
            IMPORT :: C_INT, C_PTR
            TYPE(C_PTR), VALUE :: ctx, rhs, phi
            INTEGER(C_INT) :: rc
        END FUNCTION

        FUNCTION chs_cfl_tg(ctx, dt, cflmm) BIND(C, NAME='chs_cfl_tg') RESULT(rc)
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: ctx
            REAL(C_DOUBLE), VALUE :: dt
            REAL(C_DOUBLE), INTENT(OUT) :: cflmm
            INTEGER(C_INT) :: rc
        END FUNCTION

        FUNCTION chs_max_div(ctx, maxdiv) BIND(C, NAME='chs_max_div') RESULT(rc)
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: ctx
            REAL(C_DOUBLE), INTENT(OUT) :: maxdiv
            INTEGER(C_INT) :: rc
        END FUNCTION

        FUNCTION chs_sync(ctx) BIND(C, NAME='chs_sync') RESULT(rc)
            IMPORT :: C_INT, C_PTR
            TYPE(C_PTR), VALUE :: ctx
            INTEGER(C_INT) :: rc
        END FUNCTION

        FUNCTION c_strlen(s) BIND(C, NAME='strlen') RESULT(n)
            IMPORT :: C_PTR, C_SIZE_T
            TYPE(C_PTR), VALUE :: s
            INTEGER(C_SIZE_T) :: n
        END FUNCTION
    END INTERFACE

CONTAINS

    ! Fortran string copy of the library's last error message for ctx.
    FUNCTION chs_error_message(ctx) RESULT(msg)
        TYPE(C_PTR), INTENT(IN) :: ctx
        CHARACTER(LEN=256) :: msg
        TYPE(C_PTR) :: p
        CHARACTER(KIND=C_CHAR), POINTER :: s(:)
        INTEGER :: n, i
        msg = ' '
        p = chs_last_error(ctx)
        IF (.NOT. C_ASSOCIATED(p)) RETURN
        n = INT(c_strlen(p))
        CALL C_F_POINTER(p, s, [n])
        DO i = 1, MIN(n, LEN(msg))
            msg(i:i) = s(i)
        END DO
    END FUNCTION

END MODULE chapsim_hip
