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

    ! mirror of chs_prop_table (pointers to the thermal_info arrays LIST_* and SplineCoeff_*)
    TYPE, BIND(C) :: chs_prop_table
        INTEGER(C_INT) :: n, imin_dh, imax_dh, imin_h, imax_h
        TYPE(C_PTR) :: dh, h, t, d, m, k, cp, b
        TYPE(C_PTR) :: dhh_b, dhh_c, dhh_d, ht_b, ht_c, ht_d, hd_b, hd_c, hd_d
        TYPE(C_PTR) :: hm_b, hm_c, hm_d, hk_b, hk_c, hk_d, hcp_b, hcp_c, hcp_d
        TYPE(C_PTR) :: hb_b, hb_c, hb_d
    END TYPE chs_prop_table

    ! mirror of chs_eng_params (wall_state(1:5, w) = T, H, D, M, K of wall w = 1 bottom, 2 top)
    TYPE, BIND(C) :: chs_eng_params
        INTEGER(C_INT) :: wall_type(2)
        REAL(C_DOUBLE) :: cthecd, dh_min, dh_max
        TYPE(C_PTR) :: wallflux
        REAL(C_DOUBLE) :: wall_state(5, 2)
        INTEGER(C_INT) :: top_face
    END TYPE chs_eng_params

    ! mirror of chs_io_params (gravity = |iGravity|, vis_scheme = iVisScheme)
    TYPE, BIND(C) :: chs_io_params
        REAL(C_DOUBLE) :: cvisc
        INTEGER(C_INT) :: gravity
        REAL(C_DOUBLE) :: f_a
        INTEGER(C_INT) :: flow_driven
        REAL(C_DOUBLE) :: cf_given
        INTEGER(C_INT) :: vis_scheme
    END TYPE chs_io_params

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

        FUNCTION chs_set_refinement(ctx, nref_step, nref_projection) BIND(C, NAME='chs_set_refinement') RESULT(rc)
            IMPORT :: C_INT, C_PTR
            TYPE(C_PTR), VALUE :: ctx
            INTEGER(C_INT), VALUE :: nref_step, nref_projection
            INTEGER(C_INT) :: rc
        END FUNCTION

        ! key: a C string, e.g. 'conv_by' // C_NULL_CHAR
        FUNCTION chs_set_tuning(ctx, key, val) BIND(C, NAME='chs_set_tuning') RESULT(rc)
            IMPORT :: C_INT, C_PTR, C_CHAR
            TYPE(C_PTR), VALUE :: ctx
            CHARACTER(KIND=C_CHAR), INTENT(IN) :: key(*)
            INTEGER(C_INT), VALUE :: val
            INTEGER(C_INT) :: rc
        END FUNCTION

        ! FISHPACK_POIS3D_SIMPLE (possion3d_fishpack.f90:342-453): host rhs(NCL1, NCL2, NCL3) ->
        ! phi(NCL1, NCL2, NCL3)
        FUNCTION chs_poisson_fishpack(ctx, rhs, phi) BIND(C, NAME='chs_poisson_fishpack') RESULT(rc)
            IMPORT :: C_INT, C_PTR
            TYPE(C_PTR), VALUE :: ctx, rhs, phi
            INTEGER(C_INT) :: rc
        END FUNCTION

        ! TDMAIJI_CYC (kind 0), TDMAIJJ_nonCYC (1), TDMAIJI_nonCYC (2), TDMAIJJ_CYC (3) on device
        ! arrays A, B, C, R(isz, jsz); bc(lines, 2) device boundary values of the non-cyclic kinds
        FUNCTION chs_tdma_lines(kind, isz, jsz, a, b, c, r, bc) BIND(C, NAME='chs_tdma_lines') RESULT(rc)
            IMPORT :: C_INT, C_PTR
            INTEGER(C_INT), VALUE :: kind, isz, jsz
            TYPE(C_PTR), VALUE :: a, b, c, r, bc
            INTEGER(C_INT) :: rc
        END FUNCTION

        FUNCTION chs_prop_create(t, prop) BIND(C, NAME='chs_prop_create') RESULT(rc)
            IMPORT :: C_INT, C_PTR, chs_prop_table
            TYPE(chs_prop_table), INTENT(IN) :: t
            TYPE(C_PTR), INTENT(OUT) :: prop
            INTEGER(C_INT) :: rc
        END FUNCTION

        SUBROUTINE chs_prop_destroy(prop) BIND(C, NAME='chs_prop_destroy')
            IMPORT :: C_PTR
            TYPE(C_PTR), VALUE :: prop
        END SUBROUTINE

        ! THERM_PROP_UPDATE_FROM_DH for n values: out(n, 7) = H, T, D, M, K, Cp, B
        FUNCTION chs_prop_update_from_dh(prop, n, dh, out) BIND(C, NAME='chs_prop_update_from_dh') RESULT(rc)
            IMPORT :: C_INT, C_PTR, C_LONG, C_DOUBLE
            TYPE(C_PTR), VALUE :: prop
            INTEGER(C_LONG), VALUE :: n
            REAL(C_DOUBLE), INTENT(IN) :: dh(n)
            REAL(C_DOUBLE), INTENT(OUT) :: out(n, 7)
            INTEGER(C_INT) :: rc
        END FUNCTION

        ! Energy substep SOLVERRK3_ENG_io (SOLVERRK3_ENG_IO.f90:11-82), table properties
        FUNCTION chs_eng_create(c, t, p, eng) BIND(C, NAME='chs_eng_create') RESULT(rc)
            IMPORT :: C_INT, C_PTR, chs_coeffs, chs_prop_table, chs_eng_params
            TYPE(chs_coeffs), INTENT(IN) :: c
            TYPE(chs_prop_table), INTENT(IN) :: t
            TYPE(chs_eng_params), INTENT(IN) :: p
            TYPE(C_PTR), INTENT(OUT) :: eng
            INTEGER(C_INT) :: rc
        END FUNCTION

        SUBROUTINE chs_eng_destroy(eng) BIND(C, NAME='chs_eng_destroy')
            IMPORT :: C_PTR
            TYPE(C_PTR), VALUE :: eng
        END SUBROUTINE

        ! dh = C_LOC(DH(1,0,1)), g = C_LOC(G_io(1,0,1,1)), rhs0 = C_LOC(RHS_ENERGY0) or C_NULL_PTR
        FUNCTION chs_eng_upload(eng, dh, g, rhs0, init_props) BIND(C, NAME='chs_eng_upload') RESULT(rc)
            IMPORT :: C_INT, C_PTR
            TYPE(C_PTR), VALUE :: eng, dh, g, rhs0
            INTEGER(C_INT), VALUE :: init_props
            INTEGER(C_INT) :: rc
        END FUNCTION

        FUNCTION chs_substep_eng(eng, ns, dt) BIND(C, NAME='chs_substep_eng') RESULT(rc)
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: eng
            INTEGER(C_INT), VALUE :: ns
            REAL(C_DOUBLE), VALUE :: dt
            INTEGER(C_INT) :: rc
        END FUNCTION

        FUNCTION chs_eng_download(eng, dh, h, t, d, m, k, cp, rhs0) BIND(C, NAME='chs_eng_download') RESULT(rc)
            IMPORT :: C_INT, C_PTR
            TYPE(C_PTR), VALUE :: eng, dh, h, t, d, m, k, cp, rhs0
            INTEGER(C_INT) :: rc
        END FUNCTION

        ! Main-domain momentum SOLVERRK3_MOM_io (SOLVERRK3_MOM_IO.f90:11-120) and what SOLVE runs
        ! before it (SOLVE.f90:122-127); pointers: C_LOC(G_io(1,0,1,1)), C_LOC(PR_io(1,0,1)), ...
        FUNCTION chs_io_create(c, p, io) BIND(C, NAME='chs_io_create') RESULT(rc)
            IMPORT :: C_INT, C_PTR, chs_coeffs, chs_io_params
            TYPE(chs_coeffs), INTENT(IN) :: c
            TYPE(chs_io_params), INTENT(IN) :: p
            TYPE(C_PTR), INTENT(OUT) :: io
            INTEGER(C_INT) :: rc
        END FUNCTION

        SUBROUTINE chs_io_destroy(io) BIND(C, NAME='chs_io_destroy')
            IMPORT :: C_PTR
            TYPE(C_PTR), VALUE :: io
        END SUBROUTINE

        FUNCTION chs_io_upload(io, g, g0, pr, d0, m0, explt0) BIND(C, NAME='chs_io_upload') RESULT(rc)
            IMPORT :: C_INT, C_PTR
            TYPE(C_PTR), VALUE :: io, g, g0, pr, d0, m0, explt0
            INTEGER(C_INT) :: rc
        END FUNCTION

        FUNCTION chs_io_upload_viscosity(io, visc) BIND(C, NAME='chs_io_upload_viscosity') RESULT(rc)
            IMPORT :: C_INT, C_PTR
            TYPE(C_PTR), VALUE :: io, visc
            INTEGER(C_INT) :: rc
        END FUNCTION

        ! pressure solver of the substep: 0 FFT99's gauge, 1 FISHPACK_POIS3D_SIMPLE's
        ! (is_FFT_FFT99 / is_FFT_FISHPACK, modules.f90:49-50)
        FUNCTION chs_io_set_poisson(io, fishpack) BIND(C, NAME='chs_io_set_poisson') RESULT(rc)
            IMPORT :: C_INT, C_PTR
            TYPE(C_PTR), VALUE :: io
            INTEGER(C_INT), VALUE :: fishpack
            INTEGER(C_INT) :: rc
        END FUNCTION

        ! mixed form: EXPLT0_io and RHS_io in FP32 (1) or FP64 throughout (0)
        FUNCTION chs_io_set_mixed(io, mixed) BIND(C, NAME='chs_io_set_mixed') RESULT(rc)
            IMPORT :: C_INT, C_PTR
            TYPE(C_PTR), VALUE :: io
            INTEGER(C_INT), VALUE :: mixed
            INTEGER(C_INT) :: rc
        END FUNCTION

        FUNCTION chs_io_props(io, ns) BIND(C, NAME='chs_io_props') RESULT(rc)
            IMPORT :: C_INT, C_PTR
            TYPE(C_PTR), VALUE :: io
            INTEGER(C_INT), VALUE :: ns
            INTEGER(C_INT) :: rc
        END FUNCTION

        FUNCTION chs_substep_io(io, ns, dt) BIND(C, NAME='chs_substep_io') RESULT(rc)
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: io
            INTEGER(C_INT), VALUE :: ns
            REAL(C_DOUBLE), VALUE :: dt
            INTEGER(C_INT) :: rc
        END FUNCTION

        FUNCTION chs_io_download(io, g, q, pr, dph, explt0, drho, rhsp, dstg, mstg) &
            BIND(C, NAME='chs_io_download') RESULT(rc)
            IMPORT :: C_INT, C_PTR
            TYPE(C_PTR), VALUE :: io, g, q, pr, dph, explt0, drho, rhsp, dstg, mstg
            INTEGER(C_INT) :: rc
        END FUNCTION

        ! CFL_io (CFL.f90:59-107) x dt; dt = 1 gives CFLMM_io
        FUNCTION chs_cfl_io(io, dt, cfl) BIND(C, NAME='chs_cfl_io') RESULT(rc)
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: io
            REAL(C_DOUBLE), VALUE :: dt
            REAL(C_DOUBLE), INTENT(OUT) :: cfl
            INTEGER(C_INT) :: rc
        END FUNCTION

        ! DIVGCK_io (DIVGCK.f90:60-175): MAXDIVGV_io
        FUNCTION chs_max_div_io(io, maxdiv) BIND(C, NAME='chs_max_div_io') RESULT(rc)
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: io
            REAL(C_DOUBLE), INTENT(OUT) :: maxdiv
            INTEGER(C_INT) :: rc
        END FUNCTION

        FUNCTION chs_heated_cfl(eng, io, dt, cfl) BIND(C, NAME='chs_heated_cfl') RESULT(rc)
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: eng, io
            REAL(C_DOUBLE), VALUE :: dt
            REAL(C_DOUBLE), INTENT(OUT) :: cfl
            INTEGER(C_INT) :: rc
        END FUNCTION

        FUNCTION chs_heated_substep(eng, io, ns, dt) BIND(C, NAME='chs_heated_substep') RESULT(rc)
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: eng, io
            INTEGER(C_INT), VALUE :: ns
            REAL(C_DOUBLE), VALUE :: dt
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
