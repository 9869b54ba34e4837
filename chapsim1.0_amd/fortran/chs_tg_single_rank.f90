!-----------------------------------------------------------------------------------------------
! Single-rank, GPU-resident host glue for the reference program's turbulence-generator channel.
!
! The reference's own objects are linked unchanged (DNS_THERMAL main program, READINI, SOLVE,
! the post-processing); the linker's --wrap redirects the calls below to these routines, so
! SOLVE's call structure (SOLVE.f90:74-145) stays as it is and the fields live on the GPU:
!
!   SOLVERRK3_MOM_tg(NS)   the substep on the GPU (chs_substep_tg); the host arrays Q_tg, PR_tg,
!                          CONVH0_tg are uploaded only when the host holds a newer state (the
!                          initial or restarted field)
!   CFL_tg, DIVGCK_tg      CFLMM_tg and MAXDIVGV_tg from the device fields (chs_cfl_tg,
!                          chs_max_div; the same per-cell arithmetic as CFL.f90:13-57 and
!                          DIVGCK.f90:15-58)
!   POSTPROCESS_TG         fields copied back to the host arrays only on the steps where the
!                          reference's routine reads them (averaging, monitor or save steps,
!                          the conditions of POSTPROESS.f90:17-47), then the reference's routine
!   CALL_TEC360            fields copied back, then the reference's routine
!   INTFC_VARS3/1          the y-halo exchange at one rank: the periodic rank ring of
!                          INTFC_VARS.f90:15-180 closes on the rank itself, so ghost row 0 takes
!                          row N2DO and row N2DO+1 takes row 1 (the wall routines then overwrite
!                          the wall rows, as after the multi-rank exchange).  The reference's
!                          exchange pairs each rank with its neighbours by MPI_SENDRECV and waits
!                          forever on one rank (rank -1 as the lower neighbour).
!
! So the program runs as one process on one GPU without an MPI launcher; MPI calls left in the
! reference (BCAST, ALLREDUCE, BARRIER) are single-rank no-ops.  QTMP_tg, DPH_tg, RHS_tg and
! RHSLLPHI_tg are work arrays of the substep and are not updated.  Errors stop through the
! reference's ERRHDL (WRTHDL.f90:48-62).
!-----------------------------------------------------------------------------------------------
MODULE chs_tg_resident
    USE chapsim_hip
    IMPLICIT NONE
    TYPE(C_PTR), SAVE :: ctx = C_NULL_PTR
    ! where the newest state is: 0 host arrays (upload before the next GPU use), 1 device
    ! (download before the host reads), 2 both
    INTEGER, SAVE :: newest = 0
    REAL(C_DOUBLE), ALLOCATABLE, TARGET, SAVE :: c_dyfi(:), c_dyci(:), c_wff(:), c_wfb(:)
    REAL(C_DOUBLE), ALLOCATABLE, TARGET, SAVE :: c_amvr(:, :), c_acvr(:, :), c_apvr(:, :)
    REAL(C_DOUBLE), ALLOCATABLE, TARGET, SAVE :: c_amph(:), c_acph(:), c_apph(:), c_ak1(:), c_ak3(:)
    REAL(C_DOUBLE), TARGET, SAVE :: c_tgam(0:3), c_troh(0:3), c_talp(0:3)
    ! the field copies of include/chapsim_hip.h with the reference's arrays passed by address
    INTERFACE
        FUNCTION chs_upload_arrays(c, q, pr, convh0) BIND(C, NAME='chs_upload_fields') RESULT(rc)
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: c
            REAL(C_DOUBLE) :: q(*), pr(*), convh0(*)
            INTEGER(C_INT) :: rc
        END FUNCTION
        FUNCTION chs_download_arrays(c, q, pr, convh0) BIND(C, NAME='chs_download_fields') RESULT(rc)
            IMPORT :: C_INT, C_PTR, C_DOUBLE
            TYPE(C_PTR), VALUE :: c
            REAL(C_DOUBLE) :: q(*), pr(*), convh0(*)
            INTEGER(C_INT) :: rc
        END FUNCTION
    END INTERFACE
CONTAINS
    SUBROUTINE chs_res_check(rc)
        USE mpi_info, ONLY: MYID
        INTEGER(C_INT), INTENT(IN) :: rc
        IF (rc /= CHS_OK) CALL ERRHDL('chapsim_hip: ' // TRIM(chs_error_message(ctx)), MYID)
    END SUBROUTINE chs_res_check

    ! the context from the reference's own set-up arrays (COORDJR, TDMA_COEF, FFT99_ROOT,
    ! CONS_RKCOEF); one rank holds the whole mesh
    SUBROUTINE chs_res_init
        USE mpi_info, ONLY: MYID, NPSLV
        USE mesh_info
        USE init_info
        USE FFT99_info1, ONLY: AK1, AK3
        TYPE(chs_coeffs), TARGET, SAVE :: cf
        INTEGER(C_INT) :: rc

        IF (C_ASSOCIATED(ctx)) RETURN
        IF (NPSLV /= 0) CALL ERRHDL('chapsim_hip: this program runs on one rank', MYID)
        ALLOCATE(c_dyfi(NCL2), c_dyci(NND2), c_wff(NND2), c_wfb(NND2))
        ALLOCATE(c_amvr(NND2, 3), c_acvr(NND2, 3), c_apvr(NND2, 3))
        ALLOCATE(c_amph(NCL2), c_acph(NCL2), c_apph(NCL2), c_ak1(NCL1_tg), c_ak3(NCL3))
        c_dyfi = DYFI(1:NCL2)
        c_dyci = DYCI(1:NND2)
        c_wff = YCL2ND_WFF(1:NND2)
        c_wfb = YCL2ND_WFB(1:NND2)
        c_amvr = AMVR(1:NND2, 1:3)
        c_acvr = ACVR(1:NND2, 1:3)
        c_apvr = APVR(1:NND2, 1:3)
        c_amph = AMPH(1:NCL2)
        c_acph = ACPH(1:NCL2)
        c_apph = APPH(1:NCL2)
        c_ak1 = AK1(1:NCL1_tg)
        c_ak3 = AK3(1:NCL3)
        c_tgam = TGAM
        c_troh = TROH
        c_talp = TALP
        cf%L = NCL1_tg
        cf%M = NCL2
        cf%N = NCL3
        cf%dxi = DXI
        cf%dzi = DZI
        cf%dxqi = DXQI
        cf%dzqi = DZQI
        cf%volm = VOLM_tg
        cf%dyfi = C_LOC(c_dyfi)
        cf%dyci = C_LOC(c_dyci)
        cf%wff = C_LOC(c_wff)
        cf%wfb = C_LOC(c_wfb)
        cf%amvr = C_LOC(c_amvr)
        cf%acvr = C_LOC(c_acvr)
        cf%apvr = C_LOC(c_apvr)
        cf%amph = C_LOC(c_amph)
        cf%acph = C_LOC(c_acph)
        cf%apph = C_LOC(c_apph)
        cf%ak1 = C_LOC(c_ak1)
        cf%ak3 = C_LOC(c_ak3)
        cf%tgam = C_LOC(c_tgam)
        cf%troh = C_LOC(c_troh)
        cf%talp = C_LOC(c_talp)
        rc = chs_create(cf, CHS_PLAIN_FP64, ctx)
        IF (rc /= CHS_OK) CALL ERRHDL('chapsim_hip: ' // TRIM(chs_error_message(C_NULL_PTR)), MYID)
        newest = 0
    END SUBROUTINE chs_res_init

    ! make the device hold the newest state
    SUBROUTINE chs_res_to_device
        USE flow_info, ONLY: Q_tg, PR_tg, CONVH0_tg
        CALL chs_res_init
        IF (newest == 0) THEN
            CALL chs_res_check(chs_upload_arrays(ctx, Q_tg, PR_tg, CONVH0_tg))
            newest = 2
        END IF
    END SUBROUTINE chs_res_to_device

    ! make the host arrays hold the newest state
    SUBROUTINE chs_res_to_host
        USE flow_info, ONLY: Q_tg, PR_tg, CONVH0_tg
        IF (newest == 1) THEN
            CALL chs_res_check(chs_download_arrays(ctx, Q_tg, PR_tg, CONVH0_tg))
            newest = 2
        END IF
    END SUBROUTINE chs_res_to_host
END MODULE chs_tg_resident

SUBROUTINE chs_wrap_solverrk3_mom_tg(NS) BIND(C, NAME='__wrap_solverrk3_mom_tg_')
    USE chs_tg_resident
    USE init_info, ONLY: DT, CVISC
    IMPLICIT NONE
    INTEGER(C_INT), INTENT(IN) :: NS
    CALL chs_res_to_device
    CALL chs_res_check(chs_substep_tg(ctx, NS, DT, CVISC))
    newest = 1
END SUBROUTINE chs_wrap_solverrk3_mom_tg

SUBROUTINE chs_wrap_cfl_tg() BIND(C, NAME='__wrap_cfl_tg_')
    USE chs_tg_resident
    USE flow_info, ONLY: CFLMM_tg
    IMPLICIT NONE
    REAL(C_DOUBLE) :: v
    CALL chs_res_to_device
    CALL chs_res_check(chs_cfl_tg(ctx, 1.0_C_DOUBLE, v))
    CFLMM_tg = v
END SUBROUTINE chs_wrap_cfl_tg

SUBROUTINE chs_wrap_divgck_tg() BIND(C, NAME='__wrap_divgck_tg_')
    USE chs_tg_resident
    USE flow_info, ONLY: MAXDIVGV_tg
    IMPLICIT NONE
    REAL(C_DOUBLE) :: v
    CALL chs_res_to_device
    CALL chs_res_check(chs_max_div(ctx, v))
    MAXDIVGV_tg(1) = MAXDIVGV_tg(2)
    MAXDIVGV_tg(2) = v
END SUBROUTINE chs_wrap_divgck_tg

SUBROUTINE chs_wrap_postprocess_tg() BIND(C, NAME='__wrap_postprocess_tg_')
    USE chs_tg_resident
    USE init_info, ONLY: PhyTIME, ITERG, DT, tRunAve1, iterMonitor, dtSave
    IMPLICIT NONE
    INTERFACE
        SUBROUTINE real_postprocess_tg() BIND(C, NAME='__real_postprocess_tg_')
        END SUBROUTINE real_postprocess_tg
    END INTERFACE
    ! the steps on which POSTPROCESS_TG reads the fields (POSTPROESS.f90:22-47)
    IF (PhyTIME > tRunAve1 .OR. ITERG == 1 .OR. MOD(ITERG, iterMonitor) == 0 .OR. &
        DMOD(PhyTIME, dtSave) < DT) CALL chs_res_to_host
    CALL real_postprocess_tg()
END SUBROUTINE chs_wrap_postprocess_tg

SUBROUTINE chs_wrap_call_tec360() BIND(C, NAME='__wrap_call_tec360_')
    USE chs_tg_resident
    IMPLICIT NONE
    INTERFACE
        SUBROUTINE real_call_tec360() BIND(C, NAME='__real_call_tec360_')
        END SUBROUTINE real_call_tec360
    END INTERFACE
    IF (C_ASSOCIATED(ctx)) CALL chs_res_to_host
    CALL real_call_tec360()
END SUBROUTINE chs_wrap_call_tec360

SUBROUTINE chs_wrap_intfc_vars3(L1, L2, NCL1ST, NCL1ED, VARS3) BIND(C, NAME='__wrap_intfc_vars3_')
    USE ISO_C_BINDING
    USE mpi_info, ONLY: MYID
    USE mesh_info, ONLY: N2DO, NCL3
    USE cparam, ONLY: NDV
    IMPLICIT NONE
    INTEGER(C_INT), INTENT(IN) :: L1, L2, NCL1ST, NCL1ED
    REAL(C_DOUBLE), INTENT(INOUT) :: VARS3(NCL1ST:NCL1ED, 0:N2DO(MYID) + 1, NCL3, NDV)
    INTEGER :: n2
    n2 = N2DO(MYID)
    VARS3(L1:L2, 0, :, :) = VARS3(L1:L2, n2, :, :)
    VARS3(L1:L2, n2 + 1, :, :) = VARS3(L1:L2, 1, :, :)
END SUBROUTINE chs_wrap_intfc_vars3

SUBROUTINE chs_wrap_intfc_vars1(L1, L2, NCL1ST, NCL1ED, VARS1) BIND(C, NAME='__wrap_intfc_vars1_')
    USE ISO_C_BINDING
    USE mpi_info, ONLY: MYID
    USE mesh_info, ONLY: N2DO, NCL3
    IMPLICIT NONE
    INTEGER(C_INT), INTENT(IN) :: L1, L2, NCL1ST, NCL1ED
    REAL(C_DOUBLE), INTENT(INOUT) :: VARS1(NCL1ST:NCL1ED, 0:N2DO(MYID) + 1, NCL3)
    INTEGER :: n2
    n2 = N2DO(MYID)
    VARS1(L1:L2, 0, :) = VARS1(L1:L2, n2, :)
    VARS1(L1:L2, n2 + 1, :) = VARS1(L1:L2, 1, :)
END SUBROUTINE chs_wrap_intfc_vars1
