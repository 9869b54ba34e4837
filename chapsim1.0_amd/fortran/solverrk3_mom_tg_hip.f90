!-----------------------------------------------------------------------------------------------
! Drop-in replacement for the reference's SOLVERRK3_MOM_tg(NS) (SOLVERRK3_MOM_tg.f90:13-56):
! same name, same argument, same module-global inputs and outputs, with the substep computed
! on the GPU by libchapsim_hip.so through the chapsim_hip module.
!
! Link it in place of SOLVERRK3_MOM_tg.o.  The reference decomposes y over MPI ranks
! (mesh_decomp.f90:17-64) and the GPU holds one whole channel, so each call gathers the
! velocity, pressure and RK history slabs to rank 0, runs the substep there, and scatters the
! rows back with their y-halos (the rows the reference's INTFC_VARS3/INTFC_VARS1 + BC_WALL_*
! calls at the end of the substep would produce).  QTMP_tg, DPH_tg, RHS_tg and RHSLLPHI_tg are
! work arrays of the substep and are not updated.
!
! Errors stop through the reference's ERRHDL (WRTHDL.f90:48-62), as its own routines do.
!-----------------------------------------------------------------------------------------------
MODULE chs_tg_glue
    USE chapsim_hip
    IMPLICIT NONE
    TYPE(C_PTR), SAVE :: ctx = C_NULL_PTR
    REAL(C_DOUBLE), ALLOCATABLE, TARGET, SAVE :: GQ(:, :, :, :), GP(:, :, :), GH(:, :, :, :)
    REAL(C_DOUBLE), ALLOCATABLE, TARGET, SAVE :: c_dyfi(:), c_dyci(:), c_wff(:), c_wfb(:)
    REAL(C_DOUBLE), ALLOCATABLE, TARGET, SAVE :: c_amvr(:, :), c_acvr(:, :), c_apvr(:, :)
    REAL(C_DOUBLE), ALLOCATABLE, TARGET, SAVE :: c_amph(:), c_acph(:), c_apph(:), c_ak1(:), c_ak3(:)
    REAL(C_DOUBLE), TARGET, SAVE :: c_tgam(0:3), c_troh(0:3), c_talp(0:3)
END MODULE chs_tg_glue

SUBROUTINE CHS_TG_INIT
    USE chs_tg_glue
    USE mesh_info
    USE init_info
    USE FFT99_info1, ONLY: AK1, AK3
    IMPLICIT NONE
    TYPE(chs_coeffs), TARGET, SAVE :: cf
    INTEGER(C_INT) :: rc

    ALLOCATE(GQ(NCL1_tg, 0:NCL2 + 1, NCL3, 3), GP(NCL1_tg, 0:NCL2 + 1, NCL3), GH(NCL1_tg, NCL2, NCL3, 3))
    IF (MYID /= 0) RETURN
    ! the reference's own set-up arrays (COORDJR, TDMA_COEF, FFT99_ROOT, CONS_RKCOEF)
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
END SUBROUTINE CHS_TG_INIT

! Rank r's interior rows J = JDSWT(r)..JDEWT(r) of Q_tg, PR_tg, CONVH0_tg -> GQ, GP, GH on
! rank 0, plus the wall ghost rows J = 0 (rank 0) and J = NCL2+1 (last rank).
SUBROUTINE CHS_TG_GATHER
    USE chs_tg_glue
    USE WPRECISION
    USE mpi_info, ONLY: MYID, NPSLV, ICOMM, MPI_DOUBLE_PRECISION, MPI_STATUS_SIZE
    USE mesh_info, ONLY: NCL1_tg, NCL2, NCL3, N2DO, JDSWT
    USE flow_info, ONLY: Q_tg, PR_tg, CONVH0_tg
    IMPLICIT NONE
    REAL(WP), ALLOCATABLE :: buf(:, :, :, :)
    INTEGER(4) :: ip, n2, j0, ierr, st(MPI_STATUS_SIZE), cnt

    IF (MYID /= 0) THEN
        n2 = N2DO(MYID)
        ALLOCATE(buf(NCL1_tg, n2 + 2, NCL3, 7))
        buf(:, :, :, 1:3) = Q_tg(:, 0:n2 + 1, :, 1:3)
        buf(:, :, :, 4) = PR_tg(:, 0:n2 + 1, :)
        buf(:, 2:n2 + 1, :, 5:7) = CONVH0_tg(:, 1:n2, :, 1:3)
        cnt = INT(SIZE(buf), 4)
        CALL MPI_SEND(buf, cnt, MPI_DOUBLE_PRECISION, 0, 71, ICOMM, ierr)
        DEALLOCATE(buf)
        RETURN
    END IF
    DO ip = 0, NPSLV
        n2 = N2DO(ip)
        j0 = JDSWT(ip)
        ALLOCATE(buf(NCL1_tg, n2 + 2, NCL3, 7))
        IF (ip == 0) THEN
            buf(:, :, :, 1:3) = Q_tg(:, 0:n2 + 1, :, 1:3)
            buf(:, :, :, 4) = PR_tg(:, 0:n2 + 1, :)
            buf(:, 2:n2 + 1, :, 5:7) = CONVH0_tg(:, 1:n2, :, 1:3)
        ELSE
            cnt = INT(SIZE(buf), 4)
            CALL MPI_RECV(buf, cnt, MPI_DOUBLE_PRECISION, ip, 71, ICOMM, st, ierr)
        END IF
        GQ(:, j0:j0 + n2 - 1, :, :) = buf(:, 2:n2 + 1, :, 1:3)
        GP(:, j0:j0 + n2 - 1, :) = buf(:, 2:n2 + 1, :, 4)
        GH(:, j0:j0 + n2 - 1, :, :) = buf(:, 2:n2 + 1, :, 5:7)
        IF (ip == 0) THEN
            GQ(:, 0, :, :) = buf(:, 1, :, 1:3)
            GP(:, 0, :) = buf(:, 1, :, 4)
        END IF
        IF (ip == NPSLV) THEN
            GQ(:, NCL2 + 1, :, :) = buf(:, n2 + 2, :, 1:3)
            GP(:, NCL2 + 1, :) = buf(:, n2 + 2, :, 4)
        END IF
        DEALLOCATE(buf)
    END DO
END SUBROUTINE CHS_TG_GATHER

! Rows JDSWT(r)-1 .. JDEWT(r)+1 of GQ, GP and rows JDSWT(r)..JDEWT(r) of GH back to rank r.
SUBROUTINE CHS_TG_SCATTER
    USE chs_tg_glue
    USE WPRECISION
    USE mpi_info, ONLY: MYID, NPSLV, ICOMM, MPI_DOUBLE_PRECISION, MPI_STATUS_SIZE
    USE mesh_info, ONLY: NCL1_tg, NCL2, NCL3, N2DO, JDSWT
    USE flow_info, ONLY: Q_tg, PR_tg, CONVH0_tg
    IMPLICIT NONE
    REAL(WP), ALLOCATABLE :: buf(:, :, :, :)
    INTEGER(4) :: ip, n2, j0, ierr, st(MPI_STATUS_SIZE), cnt

    IF (MYID /= 0) THEN
        n2 = N2DO(MYID)
        ALLOCATE(buf(NCL1_tg, n2 + 2, NCL3, 7))
        cnt = INT(SIZE(buf), 4)
        CALL MPI_RECV(buf, cnt, MPI_DOUBLE_PRECISION, 0, 72, ICOMM, st, ierr)
        Q_tg(:, 0:n2 + 1, :, 1:3) = buf(:, :, :, 1:3)
        PR_tg(:, 0:n2 + 1, :) = buf(:, :, :, 4)
        CONVH0_tg(:, 1:n2, :, 1:3) = buf(:, 2:n2 + 1, :, 5:7)
        DEALLOCATE(buf)
        RETURN
    END IF
    DO ip = NPSLV, 0, -1
        n2 = N2DO(ip)
        j0 = JDSWT(ip)
        ALLOCATE(buf(NCL1_tg, n2 + 2, NCL3, 7))
        buf = 0.0_WP
        buf(:, :, :, 1:3) = GQ(:, j0 - 1:j0 + n2, :, :)
        buf(:, :, :, 4) = GP(:, j0 - 1:j0 + n2, :)
        buf(:, 2:n2 + 1, :, 5:7) = GH(:, j0:j0 + n2 - 1, :, :)
        IF (ip == 0) THEN
            Q_tg(:, 0:n2 + 1, :, 1:3) = buf(:, :, :, 1:3)
            PR_tg(:, 0:n2 + 1, :) = buf(:, :, :, 4)
            CONVH0_tg(:, 1:n2, :, 1:3) = buf(:, 2:n2 + 1, :, 5:7)
        ELSE
            cnt = INT(SIZE(buf), 4)
            CALL MPI_SEND(buf, cnt, MPI_DOUBLE_PRECISION, ip, 72, ICOMM, ierr)
        END IF
        DEALLOCATE(buf)
    END DO
END SUBROUTINE CHS_TG_SCATTER

SUBROUTINE SOLVERRK3_MOM_tg(NS)
    USE chs_tg_glue
    USE cparam
    USE flow_info
    USE mesh_info
    USE init_info
    IMPLICIT NONE
    INTEGER(4), INTENT(IN) :: NS
    INTEGER(C_INT) :: rc

    IF (.NOT. ALLOCATED(GQ)) CALL CHS_TG_INIT
    CALL CHS_TG_GATHER
    IF (MYID == 0) THEN
        rc = chs_upload_fields(ctx, C_LOC(GQ), C_LOC(GP), C_LOC(GH))
        IF (rc == CHS_OK) rc = chs_substep_tg(ctx, INT(NS, C_INT), DT, CVISC)
        IF (rc == CHS_OK) rc = chs_download_fields(ctx, C_LOC(GQ), C_LOC(GP), C_LOC(GH))
        IF (rc /= CHS_OK) CALL ERRHDL('chapsim_hip: ' // TRIM(chs_error_message(ctx)), MYID)
    END IF
    CALL CHS_TG_SCATTER
    RETURN
END SUBROUTINE SOLVERRK3_MOM_tg
