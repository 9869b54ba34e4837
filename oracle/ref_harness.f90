!-----------------------------------------------------------------------------------------------
! ref_harness: test infrastructure only.  Drives single routines of the unmodified CHAPSim 1.0
! reference (linked from oracle/_ref/obj by oracle/ref.mk) on inputs given as raw files, and
! dumps what they compute, so the C restatement in oracle/ and the HIP path can be checked
! against the reference itself.
!
! Set-up repeats the reference main program's sequence (DNS_THERMAL.f90:59-130): READINI,
! decomposition, MEM_ALLOCAT, index maps, RK coefficients, mesh, Poiseuille profile, TDMA
! coefficients and FFT99_POIS3D_INIT(ITG).  Then commands are read from ./harness.cmd, one
! per line:
!   dumpsetup FILE          rank 0 writes the set-up arrays (layout in oracle/refrun.py)
!   load FILE               every rank reads global U,V,W,P (each L x (M+2) x N, y-halo
!                           rows 0..M+1, Fortran order, float64) and keeps its y-slab
!   substep NS DT           DT = given, CVISC = 1/REN, CALL SOLVERRK3_MOM_tg(NS)
!   dump FILE               each rank writes FILE.<rank>: Q_tg and PR_tg rows 0..N2DO+1
!   conv FILE               CONVECTION_X/Y/Z_tg, then dumps QTMP, DPH, RHSLLPHI rows 1..N2DO
!   divg NS DT FILE         DIVG_tg(NS) and dumps RHSLLPHI_tg rows 1..N2DO
!   poisson FILEIN FILE     RHSLLPHI_tg from a global L x M x N file, FFT99 Poisson solve,
!                           dumps DPH_tg rows 1..N2DO
!   rhsmom NS DT FILE       conv + RHS_CvLpGpS_tg/MOMFA_tg for IDR=1..3 (as SOLVERRK3_MOM_tg
!                           does), dumps Q_tg rows 0..N2DO+1 and RHS_tg rows 1..N2DO
!   thermoinit              rank 0: thermal_init (table read, sort, nondimensionalisation,
!                           splines; Flow_State.f90:16-95), needs iThermoDynamics = 1
!   thermoinitall           thermal_init on rank 0 and BCAST_THERMAL_INIT (all ranks)
!   dumpengpar FILE         rank 0: CTHECD, DHmin, DHmax, WALLFLUX(1, bottom/top, 1), the wall
!                           states T, H, D, M, K (bottom, top) and N2DO of the last rank
!   dumptable FILE          rank 0: nondimensional table and spline coefficients
!   proplookup FIN FOUT     rank 0: THERM_PROP_UPDATE_FROM_DH for every DH in FIN (int32 n,
!                           then n float64); FOUT = H, T, D, M, K, Cp, B (7 x n float64)
!   poissonio FIN FOUT      (main/IO domain) RHSLLPHI_io from a global L x M x N file,
!                           FISHPACK_POIS3D_SIMPLE into an array of its declared shape,
!                           dumps the solution rows 1..N2DO
!   poissonio_call FIN FOUT same, with DPH_io as the output argument as the IO substep
!                           passes it; dumps DPH_io rows 1..N2DO
!   loadeng FILE            (main domain, iThermoDynamics = 1) DH and G_io from FILE, see
!                           H_LOADENG
!   engsubstep NS DT        DT = given, CALL SOLVERRK3_ENG_io(NS)
!   dumpeng FILE            each rank writes FILE.<rank>: DH, H, T, D, M, K, Cp (x 1..L, rows
!                           0..N2DO+1), then RHS_ENERGY0 rows 1..N2DO
!   loadio FILE             (main domain, iThermoDynamics = 1) G_io, G0_io, PR_io, DENSITY0,
!                           Viscousity0 and EXPLT0_io from FILE, see H_LOADIO
!   ioprops NS              DENSITY_TIME_DERIVATION(NS), DENSITY_Staggered, MU_Staggered
!                           and, with iVisScheme = 1, DENSITY_IMPLICIT_add (SOLVE.f90:122-127)
!   loadvisc FILE           Viscousity (the current viscosity) from a global L x (M+2) x N file
!   cfltg FILE              CFL_tg; rank 0 writes CFLMM_tg to FILE
!   divcktg FILE            DIVGCK_tg; rank 0 writes MAXDIVGV_tg(2) to FILE
!   cflio FILE              VELOCITY_CALC_io and CFL_io; rank 0 writes CFLMM_io to FILE
!   divckio FILE            DIVGCK_io; rank 0 writes MAXDIVGV_io(1:3) to FILE
!   iocn NS DT FILE         DT = given; the momentum part of SOLVERRK3_MOM_io(NS) with the
!                           Crank-Nicolson viscous scheme (SOLVERRK3_MOM_IO.f90:25-41, :61-74):
!                           each rank writes FILE.<rank>: G_io and DRHOI_STG rows 0..N2DO+1
!                           after the three MOMFA_io calls, then RHS_io of each component as
!                           MOMFA_io receives it and EXPLT0_io, rows 1..N2DO
!   ioterms FILE            VELOCITY_CALC_io, CONVECTION_X/Y/Z_io, DIVG_U_io and
!                           VISCOUS_ALL_EXPLT_X/Y/Z_io; each rank writes FILE.<rank>: QTMP_io,
!                           DPH_io, RHSLLPHI_io, DivU_io rows 1..N2DO
!   iosubstep NS DT         DT = given; the body of SOLVERRK3_MOM_io(NS) (either viscous
!                           scheme) with the FFT99 Poisson solve, see H_IOSUBSTEP
!   dumpio FILE             each rank writes FILE.<rank>: G_io, Q_io, PR_io, D_STG, MU_STG,
!                           DPH_io rows 0..N2DO+1, then EXPLT0_io, DrhoDtP, RHSLLPHI_io rows
!                           1..N2DO
!   dumpiopar FILE          rank 0: F_A, VOLM_io, CVISC
!   meanzx FILE             PP_MEAN_ZX_FLOW_TG on the loaded TG field; each rank writes
!                           FILE.<rank>: U1xzL_tg, UPxzL_tg, U2xzL_tg, U3xzL_tg, DVDL1xzL_tg,
!                           DVDLPxzL_tg, DVDL2xzL_tg (rows 1..N2DO, Fortran order), then
!                           U1mean_WORK_tg, U1maxx_WORK_tg
!   rstwrite T IT           PhyTIME = T, ITERG = IT, WRT_INSTANT_VARS_TG (into FilePath1)
!   rstread T               ReStart_INSTANT_VARS_TG(T) and the walls as FLOWSTART.f90:37-40
!   rstwriteio T IT         PhyTIME = T, ITERG = IT, VELOCITY_CALC_io, WRT_INSTANT_VARS_IO
!   rstreadio T             ReStart_INSTANT_VARS_io(T) (main domain, iIniFieldType of the ini)
!   interp                  INITIAL_INTERP_tg (needs iIniField_tg = 1, the coarse files in
!                           FilePath3 and INTP_COORD_YND.ini), then the walls as
!                           FLOWSTART.f90:37-40
!   randomtg                random_FL_FLD_TG (iRandomType of the ini; before the projection
!                           substep, which is 'substep 0 DT')
!   tdma KIND FIN FOUT      rank 0: one of the reference's batched Thomas solvers on FIN
!                           (int32 isz, jsz; A, B, C, R as isz x jsz float64, Fortran order;
!                           for the non-cyclic kinds BC as nlines x 2 float64); KIND 0
!                           TDMAIJI_CYC, 1 TDMAIJJ_nonCYC, 2 TDMAIJI_nonCYC, 3 TDMAIJJ_CYC;
!                           FOUT = R after the solve
! Must run on 2 or more MPI ranks (the reference deadlocks on one, INTFC_VARS.f90:57-79).
!-----------------------------------------------------------------------------------------------
PROGRAM ref_harness
    USE postprocess_info
    USE wrt_info
    USE init_info
    USE mesh_info
    USE flow_info
    USE thermal_info
    USE FFT99_info1
    USE FISHPACK_POIS3D
    IMPLICIT NONE

    CHARACTER(LEN=512) :: line, cmd, fname, fname2
    INTEGER(4) :: ios, ns, idr
    REAL(WP) :: dtin
    REAL(WP), ALLOCATABLE :: VIO(:, :, :)

    CALL MKDIR
    CALL Start_MPI
    IF (MYID == 0) THEN
        LOGFIL_PG = TRIM(FilePath0) // 'CHAPSim_harness.log'
        OPEN (logflg_pg, FILE = TRIM(LOGFIL_PG), action = 'write')
    END IF
    IF (MYID == 0) CALL READINI
    CALL BCAST_READINI
    CALL mesh_Ydecomp
    CALL mesh_Zdecomp
    CALL MEM_ALLOCAT
    CALL INDEXL2G
    CALL INDEXSET
    IF (MYID == 0) CALL CONS_RKCOEF
    CALL BCAST_RKCOEF
    IF (MYID == 0) CALL CONSPARA
    CALL BCAST_CONSPARA
    IF (MYID == 0) CALL COORDJR
    CALL BCAST_COORDJR
    IF (MYID == 0) CALL LAMPOISLPROF
    CALL BCAST_LAMPOISLPROF
    IF (MYID == 0) CALL TDMA_COEF
    CALL BCAST_TDMA_COEF
    IF (TgFlowFlg) CALL FFT99_POIS3D_INIT(ITG)
    IF (IoFlowFlg) CALL FISHPACK_POIS3D_INIT   ! DNS_THERMAL.f90:128 (is_FFT_FISHPACK)
    ! the main domain's other Poisson method (DNS_THERMAL.f90:127), used by iosubstep
    IF (IoFlowFlg .AND. .NOT. TgFlowFlg) CALL FFT99_POIS3D_INIT(IIO)
    CALL MPI_BARRIER(ICOMM, IERROR)

    IF (TgFlowFlg) CONVH0_tg = 0.0_WP
    CVISC = 1.0_WP / REN

    OPEN(77, FILE = 'harness.cmd', STATUS = 'old', ACTION = 'read')
    DO
        READ(77, '(A)', IOSTAT = ios) line
        IF (ios /= 0) EXIT
        IF (LEN_TRIM(line) == 0) CYCLE
        READ(line, *) cmd
        SELECT CASE (TRIM(cmd))
        CASE ('dumpsetup')
            READ(line, *) cmd, fname
            IF (MYID == 0) CALL H_DUMPSETUP(TRIM(fname))
        CASE ('load')
            READ(line, *) cmd, fname
            CALL H_LOAD(TRIM(fname))
        CASE ('substep')
            READ(line, *) cmd, ns, dtin
            DT = dtin
            CVISC = 1.0_WP / REN
            CALL SOLVERRK3_MOM_tg(ns)
        CASE ('dump')
            READ(line, *) cmd, fname
            CALL H_DUMPQ(TRIM(fname))
        CASE ('conv')
            READ(line, *) cmd, fname
            CALL CONVECTION_X_tg
            CALL CONVECTION_Y_tg
            CALL CONVECTION_Z_tg
            CALL H_DUMPCONV(TRIM(fname))
        CASE ('divg')
            READ(line, *) cmd, ns, dtin, fname
            DT = dtin
            CALL DIVG_tg(ns)
            CALL H_DUMP3(TRIM(fname), RHSLLPHI_tg, 1)
        CASE ('poisson')
            READ(line, *) cmd, fname2, fname
            CALL H_LOADRHS(TRIM(fname2))
            CALL FFT99_POIS3D_periodicxz(ITG)
            CALL H_DUMP3H(TRIM(fname), DPH_tg)
        CASE ('rhsmom')
            READ(line, *) cmd, ns, dtin, fname
            DT = dtin
            CVISC = 1.0_WP / REN
            CALL CONVECTION_X_tg
            CALL CONVECTION_Y_tg
            CALL CONVECTION_Z_tg
            DO idr = 1, 3
                CALL RHS_CvLpGpS_tg(ns, idr)
                CALL MOMFA_tg(ns, idr)
            END DO
            CALL H_DUMPQ(TRIM(fname))
            CALL H_DUMP3(TRIM(fname) // '.rhs', RHS_tg, 1)
        CASE ('thermoinit')
            IF (MYID == 0) CALL thermal_init
        CASE ('thermoinitall')
            IF (MYID == 0) CALL thermal_init
            CALL BCAST_THERMAL_INIT
        CASE ('dumpengpar')
            READ(line, *) cmd, fname
            IF (MYID == 0) THEN
                OPEN(80, FILE = TRIM(fname), ACCESS = 'STREAM', FORM = 'UNFORMATTED', &
                     STATUS = 'replace', ACTION = 'write')
                WRITE(80) CTHECD, DHmin, DHmax
                IF (ALLOCATED(WALLFLUX)) THEN
                    WRITE(80) WALLFLUX(1, iBotWall, 1), WALLFLUX(1, iTopWall, 1)
                ELSE
                    WRITE(80) 0.0_WP, 0.0_WP
                END IF
                IF (ALLOCATED(T_WAL_GV)) THEN
                    WRITE(80) T_WAL_GV(1, 1:2), H_WAL_GV(1, 1:2), D_WAL_GV(1, 1:2), M_WAL_GV(1, 1:2), &
                              K_WAL_GV(1, 1:2)
                ELSE
                    WRITE(80) (0.0_WP, idr = 1, 10)
                END IF
                WRITE(80) REAL(N2DO(NPSLV), WP)
                CLOSE(80)
            END IF
        CASE ('dumptable')
            READ(line, *) cmd, fname
            IF (MYID == 0) CALL H_DUMPTABLE(TRIM(fname))
        CASE ('proplookup')
            READ(line, *) cmd, fname2, fname
            IF (MYID == 0) CALL H_PROPLOOKUP(TRIM(fname2), TRIM(fname))
        CASE ('poissonio')
            ! the solver's output argument is VAR(NCL1_io, N2DO(0), NCL3); SOLVERRK3_MOM_IO.f90:90-93
            ! passes DPH_io, which has y-halo rows (MEMALLOCT.f90:157), so the harness gives it an
            ! array of the declared shape to pin the solver itself
            READ(line, *) cmd, fname2, fname
            CALL H_LOADRHSIO(TRIM(fname2))
            ALLOCATE(VIO(NCL1_io, N2DO(0), NCL3))
            CALL FISHPACK_POIS3D_SIMPLE(RHSLLPHI_io, VIO)
            CALL H_OPENRANK(TRIM(fname), 80)
            WRITE(80) VIO(:, 1:N2DO(MYID), :)
            CLOSE(80)
            DEALLOCATE(VIO)
        CASE ('poissonio_call')
            ! as SOLVERRK3_MOM_IO.f90:93 calls it (DPH_io as the output argument)
            READ(line, *) cmd, fname2, fname
            CALL H_LOADRHSIO(TRIM(fname2))
            CALL FISHPACK_POIS3D_SIMPLE(RHSLLPHI_io, DPH_io)
            CALL H_OPENRANK(TRIM(fname), 80)
            WRITE(80) DPH_io(1:NCL1_io, 1:N2DO(MYID), :)
            CLOSE(80)
        CASE ('loadeng')
            READ(line, *) cmd, fname
            CALL H_LOADENG(TRIM(fname))
        CASE ('engsubstep')
            READ(line, *) cmd, ns, dtin
            DT = dtin
            CALL SOLVERRK3_ENG_io(ns)
        CASE ('dumpeng')
            READ(line, *) cmd, fname
            CALL H_OPENRANK(TRIM(fname), 80)
            WRITE(80) DH(1:NCL1_io, :, :), ENTHALPY(1:NCL1_io, :, :), TEMPERATURE(1:NCL1_io, :, :), &
                      DENSITY(1:NCL1_io, :, :), Viscousity(1:NCL1_io, :, :), THERMCONDT(1:NCL1_io, :, :), &
                      HEATCAP(1:NCL1_io, :, :)
            WRITE(80) RHS_ENERGY0(:, 1:N2DO(MYID), :)
            CLOSE(80)
        CASE ('loadio')
            READ(line, *) cmd, fname
            CALL H_LOADIO(TRIM(fname))
        CASE ('ioprops')
            READ(line, *) cmd, ns
            CALL DENSITY_TIME_DERIVATION(ns)
            CALL DENSITY_Staggered
            CALL MU_Staggered
            IF (iVisScheme == VisImplicit) CALL DENSITY_IMPLICIT_add
        CASE ('loadvisc')
            READ(line, *) cmd, fname
            CALL H_LOADVISC(TRIM(fname))
        CASE ('cflio')
            READ(line, *) cmd, fname
            CALL VELOCITY_CALC_io
            CALL CFL_io
            IF (MYID == 0) THEN
                OPEN(80, FILE = TRIM(fname), ACCESS = 'STREAM', FORM = 'UNFORMATTED', &
                     STATUS = 'replace', ACTION = 'write')
                WRITE(80) CFLMM_io
                CLOSE(80)
            END IF
        CASE ('cfltg')
            READ(line, *) cmd, fname
            CALL CFL_tg
            IF (MYID == 0) THEN
                OPEN(80, FILE = TRIM(fname), ACCESS = 'STREAM', FORM = 'UNFORMATTED', &
                     STATUS = 'replace', ACTION = 'write')
                WRITE(80) CFLMM_tg
                CLOSE(80)
            END IF
        CASE ('divcktg')
            READ(line, *) cmd, fname
            CALL DIVGCK_tg
            IF (MYID == 0) THEN
                OPEN(80, FILE = TRIM(fname), ACCESS = 'STREAM', FORM = 'UNFORMATTED', &
                     STATUS = 'replace', ACTION = 'write')
                WRITE(80) MAXDIVGV_tg(2)
                CLOSE(80)
            END IF
        CASE ('divckio')
            READ(line, *) cmd, fname
            CALL DIVGCK_io
            IF (MYID == 0) THEN
                OPEN(80, FILE = TRIM(fname), ACCESS = 'STREAM', FORM = 'UNFORMATTED', &
                     STATUS = 'replace', ACTION = 'write')
                WRITE(80) MAXDIVGV_io
                CLOSE(80)
            END IF
        CASE ('iocn')
            READ(line, *) cmd, ns, dtin, fname
            DT = dtin
            CALL H_IOCN(ns, TRIM(fname))
        CASE ('ioterms')
            READ(line, *) cmd, fname
            CALL VELOCITY_CALC_io
            CALL CONVECTION_X_io
            CALL CONVECTION_Y_io
            CALL CONVECTION_Z_io
            CALL DIVG_U_io
            CALL VISCOUS_ALL_EXPLT_X_io
            CALL VISCOUS_ALL_EXPLT_Y_io
            CALL VISCOUS_ALL_EXPLT_Z_io
            CALL H_OPENRANK(TRIM(fname), 80)
            WRITE(80) QTMP_io(1:NCL1_io, 1:N2DO(MYID), :), DPH_io(1:NCL1_io, 1:N2DO(MYID), :), &
                      RHSLLPHI_io(1:NCL1_io, 1:N2DO(MYID), :), DivU_io(1:NCL1_io, 1:N2DO(MYID), :)
            CLOSE(80)
        CASE ('iosubstep')
            READ(line, *) cmd, ns, dtin
            DT = dtin
            CALL H_IOSUBSTEP(ns)
        CASE ('dumpio')
            READ(line, *) cmd, fname
            CALL H_OPENRANK(TRIM(fname), 80)
            WRITE(80) G_io(1:NCL1_io, 0:N2DO(MYID) + 1, :, :), Q_io(1:NCL1_io, 0:N2DO(MYID) + 1, :, :), &
                      PR_io(1:NCL1_io, 0:N2DO(MYID) + 1, :), D_STG(1:NCL1_io, 0:N2DO(MYID) + 1, :, :), &
                      MU_STG(1:NCL1_io, 0:N2DO(MYID) + 1, :, :), DPH_io(1:NCL1_io, 0:N2DO(MYID) + 1, :)
            WRITE(80) EXPLT0_io(:, 1:N2DO(MYID), :, :), DrhoDtP(1:NCL1_io, 1:N2DO(MYID), :), &
                      RHSLLPHI_io(1:NCL1_io, 1:N2DO(MYID), :)
            CLOSE(80)
        CASE ('dumpiopar')
            READ(line, *) cmd, fname
            IF (MYID == 0) THEN
                OPEN(80, FILE = TRIM(fname), ACCESS = 'STREAM', FORM = 'UNFORMATTED', &
                     STATUS = 'replace', ACTION = 'write')
                WRITE(80) F_A, VOLM_io, CVISC
                CLOSE(80)
            END IF
        CASE ('meanzx')
            READ(line, *) cmd, fname
            CALL PP_MEAN_ZX_FLOW_TG
            CALL H_OPENRANK(TRIM(fname), 80)
            WRITE(80) U1xzL_tg(1:N2DO(MYID), :), UPxzL_tg(1:N2DO(MYID), :), U2xzL_tg(1:N2DO(MYID), :), &
                      U3xzL_tg(1:N2DO(MYID), :), DVDL1xzL_tg(1:N2DO(MYID), :, :), &
                      DVDLPxzL_tg(1:N2DO(MYID), :, :), DVDL2xzL_tg(1:N2DO(MYID), :, :)
            WRITE(80) U1mean_WORK_tg, U1maxx_WORK_tg
            CLOSE(80)
        CASE ('rstwrite')
            READ(line, *) cmd, dtin, idr
            PhyTIME = dtin
            ITERG = idr
            CALL WRT_INSTANT_VARS_TG
        CASE ('rstwriteio')
            READ(line, *) cmd, dtin, idr
            PhyTIME = dtin
            ITERG = idr
            CALL VELOCITY_CALC_io
            CALL WRT_INSTANT_VARS_IO
        CASE ('rstreadio')
            READ(line, *) cmd, dtin
            CALL ReStart_INSTANT_VARS_io(dtin)
        CASE ('rstread')
            READ(line, *) cmd, dtin
            CALL ReStart_INSTANT_VARS_TG(dtin)
            CALL INTFC_VARS1(1, NCL1_tg, 1, NCL1_tg, PR_tg)
            CALL INTFC_VARS3(1, NCL1_tg, 1, NCL1_tg, Q_tg)
            CALL BC_WALL_Q_tg
            CALL BC_WALL_PR_TG
        CASE ('interp')
            CALL INITIAL_INTERP_tg
            CALL INTFC_VARS1(1, NCL1_tg, 1, NCL1_tg, PR_tg)
            CALL INTFC_VARS3(1, NCL1_tg, 1, NCL1_tg, Q_tg)
            CALL BC_WALL_Q_tg
            CALL BC_WALL_PR_TG
        CASE ('randomtg')
            CALL random_FL_FLD_TG
        CASE ('tdma')
            READ(line, *) cmd, idr, fname2, fname
            IF (MYID == 0) CALL H_TDMA(idr, TRIM(fname2), TRIM(fname))
        CASE DEFAULT
            IF (MYID == 0) WRITE(*, *) 'ref_harness: unknown command ', TRIM(cmd)
            CALL MPI_ABORT(ICOMM, 2, IERROR)
        END SELECT
        CALL MPI_BARRIER(ICOMM, IERROR)
    END DO
    CLOSE(77)

    CALL MPI_FINALIZE(IERROR)

CONTAINS

    SUBROUTINE H_OPENRANK(fn, unit)
        CHARACTER(*), INTENT(IN) :: fn
        INTEGER(4), INTENT(IN) :: unit
        CHARACTER(LEN=16) :: rk
        WRITE(rk, '(I0)') MYID
        OPEN(unit, FILE = fn // '.' // TRIM(rk), ACCESS = 'STREAM', FORM = 'UNFORMATTED', &
             STATUS = 'replace', ACTION = 'write')
        WRITE(unit) INT(MYID, 4), INT(JDSWT(MYID), 4), INT(N2DO(MYID), 4), &
                    INT(NCL1_tg + NCL1_io, 4), INT(NCL2, 4), INT(NCL3, 4)
    END SUBROUTINE

    SUBROUTINE H_DUMPSETUP(fn)
        CHARACTER(*), INTENT(IN) :: fn
        OPEN(78, FILE = fn, ACCESS = 'STREAM', FORM = 'UNFORMATTED', STATUS = 'replace', &
             ACTION = 'write')
        WRITE(78) INT(NCL1_tg, 4), INT(NCL2, 4), INT(NCL3, 4)
        WRITE(78) IPV_tg(1:NCL1_tg), IMV_tg(1:NCL1_tg), KPV(1:NCL3), KMV(1:NCL3), KSYM(1:NCL3)
        WRITE(78) JGPV(0:NCL2), JGMV(1:NND2)
        WRITE(78) DX, DZ, DXI, DZI, DXQI, DZQI, VOLM_tg, Area_inlet, AMVR1, APVRN
        WRITE(78) TGAM(0:3), TROH(0:3), TALP(0:3)
        WRITE(78) YND(1:NND2), YCC(1:NCL2), DYFI(1:NCL2), DYCI(1:NND2)
        WRITE(78) YCL2ND_WFF(1:NND2), YCL2ND_WFB(1:NND2)
        WRITE(78) AMVR(1:NND2, 1:3), ACVR(1:NND2, 1:3), APVR(1:NND2, 1:3)
        WRITE(78) AMPH(1:NCL2), ACPH(1:NCL2), APPH(1:NCL2)
        WRITE(78) AK1(1:NCL1_tg), AK3(1:NCL3)
        WRITE(78) Vini(1:NCL2)
        CLOSE(78)
    END SUBROUTINE

    SUBROUTINE H_LOAD(fn)
        CHARACTER(*), INTENT(IN) :: fn
        REAL(WP), ALLOCATABLE :: G(:, :, :, :)
        INTEGER(4) :: J, JG
        ALLOCATE(G(NCL1_tg, 0:NCL2 + 1, NCL3, 4))
        OPEN(79, FILE = fn, ACCESS = 'STREAM', FORM = 'UNFORMATTED', STATUS = 'old', &
             ACTION = 'read')
        READ(79) G
        CLOSE(79)
        DO J = 0, N2DO(MYID) + 1
            JG = JDSWT(MYID) - 1 + J
            Q_tg(:, J, :, 1) = G(:, JG, :, 1)
            Q_tg(:, J, :, 2) = G(:, JG, :, 2)
            Q_tg(:, J, :, 3) = G(:, JG, :, 3)
            PR_tg(:, J, :)   = G(:, JG, :, 4)
        END DO
        DEALLOCATE(G)
    END SUBROUTINE

    SUBROUTINE H_LOADRHS(fn)
        CHARACTER(*), INTENT(IN) :: fn
        REAL(WP), ALLOCATABLE :: G(:, :, :)
        INTEGER(4) :: J, JG
        ALLOCATE(G(NCL1_tg, NCL2, NCL3))
        OPEN(79, FILE = fn, ACCESS = 'STREAM', FORM = 'UNFORMATTED', STATUS = 'old', &
             ACTION = 'read')
        READ(79) G
        CLOSE(79)
        DO J = 1, N2DO(MYID)
            JG = JDSWT(MYID) - 1 + J
            RHSLLPHI_tg(:, J, :) = G(:, JG, :)
        END DO
        DEALLOCATE(G)
    END SUBROUTINE

    SUBROUTINE H_DUMPQ(fn)
        CHARACTER(*), INTENT(IN) :: fn
        INTEGER(4) :: n
        CALL H_OPENRANK(fn, 80)
        DO n = 1, 3
            WRITE(80) Q_tg(:, 0:N2DO(MYID) + 1, :, n)
        END DO
        WRITE(80) PR_tg(:, 0:N2DO(MYID) + 1, :)
        CLOSE(80)
    END SUBROUTINE

    SUBROUTINE H_DUMPCONV(fn)
        CHARACTER(*), INTENT(IN) :: fn
        CALL H_OPENRANK(fn, 80)
        WRITE(80) QTMP_tg(:, 1:N2DO(MYID), :)
        WRITE(80) DPH_tg(:, 1:N2DO(MYID), :)
        WRITE(80) RHSLLPHI_tg(:, 1:N2DO(MYID), :)
        CLOSE(80)
    END SUBROUTINE

    SUBROUTINE H_DUMP3(fn, A, j0)
        CHARACTER(*), INTENT(IN) :: fn
        INTEGER(4), INTENT(IN) :: j0
        REAL(WP), INTENT(IN) :: A(NCL1_tg, j0:N2DO(0) + j0 - 1, NCL3)
        CALL H_OPENRANK(fn, 80)
        WRITE(80) A(:, 1:N2DO(MYID), :)
        CLOSE(80)
    END SUBROUTINE

    SUBROUTINE H_LOADRHSIO(fn)
        CHARACTER(*), INTENT(IN) :: fn
        REAL(WP), ALLOCATABLE :: G(:, :, :)
        INTEGER(4) :: J, JG
        ALLOCATE(G(NCL1_io, NCL2, NCL3))
        OPEN(79, FILE = fn, ACCESS = 'STREAM', FORM = 'UNFORMATTED', STATUS = 'old', &
             ACTION = 'read')
        READ(79) G
        CLOSE(79)
        DO J = 1, N2DO(MYID)
            JG = JDSWT(MYID) - 1 + J
            RHSLLPHI_io(:, J, :) = G(:, JG, :)
        END DO
        DEALLOCATE(G)
    END SUBROUTINE
    ! Energy-equation state from one file: global DH (L x (M+2) x N, y rows 0..M+1) then G_io
    ! (same, three components).  Every rank keeps its y-slab with halo rows; properties of rows
    ! 0..N2DO+1 follow from DH (THERM_PROP_UPDATE_FROM_DH), then the walls are set as the
    ! substep sets them (BC_WALL_THERMAL(IMAIND)); RHS_ENERGY0 = 0.
    SUBROUTINE H_LOADENG(fn)
        CHARACTER(*), INTENT(IN) :: fn
        REAL(WP), ALLOCATABLE :: G(:, :, :, :)
        INTEGER(4) :: I, J, K, JG
        REAL(WP) :: B_tmp
        ALLOCATE(G(NCL1_io, 0:NCL2 + 1, NCL3, 4))
        OPEN(79, FILE = fn, ACCESS = 'STREAM', FORM = 'UNFORMATTED', STATUS = 'old', &
             ACTION = 'read')
        READ(79) G
        CLOSE(79)
        DO J = 0, N2DO(MYID) + 1
            JG = JDSWT(MYID) - 1 + J
            DH(1:NCL1_io, J, :) = G(:, JG, :, 1)
            G_io(1:NCL1_io, J, :, 1) = G(:, JG, :, 2)
            G_io(1:NCL1_io, J, :, 2) = G(:, JG, :, 3)
            G_io(1:NCL1_io, J, :, 3) = G(:, JG, :, 4)
            DO K = 1, NCL3
                DO I = 1, NCL1_io
                    CALL THERM_PROP_UPDATE_FROM_DH(DH(I, J, K), ENTHALPY(I, J, K), TEMPERATURE(I, J, K), &
                         DENSITY(I, J, K), Viscousity(I, J, K), THERMCONDT(I, J, K), HEATCAP(I, J, K), B_tmp)
                END DO
            END DO
        END DO
        CALL BC_WALL_THERMAL(IMAIND)
        RHS_ENERGY0 = 0.0_WP
        DEALLOCATE(G)
    END SUBROUTINE
    ! Main-domain momentum state from one file: 12 global arrays, each L x (M+2) x N (y rows
    ! 0..M+1, Fortran order): G_io (3 components), G0_io (3), PR_io, DENSITY0, Viscousity0,
    ! EXPLT0_io (3, rows 1..M used).  Every rank keeps its y-slab; DENSITY and Viscousity are
    ! set equal to DENSITY0 and Viscousity0.
    SUBROUTINE H_LOADIO(fn)
        CHARACTER(*), INTENT(IN) :: fn
        REAL(WP), ALLOCATABLE :: A(:, :, :, :)
        INTEGER(4) :: J, JG, C
        ALLOCATE(A(NCL1_io, 0:NCL2 + 1, NCL3, 12))
        OPEN(79, FILE = fn, ACCESS = 'STREAM', FORM = 'UNFORMATTED', STATUS = 'old', &
             ACTION = 'read')
        READ(79) A
        CLOSE(79)
        DO J = 0, N2DO(MYID) + 1
            JG = JDSWT(MYID) - 1 + J
            DO C = 1, 3
                G_io(1:NCL1_io, J, :, C) = A(:, JG, :, C)
                G0_io(1:NCL1_io, J, :, C) = A(:, JG, :, 3 + C)
            END DO
            PR_io(1:NCL1_io, J, :) = A(:, JG, :, 7)
            DENSITY0(1:NCL1_io, J, :) = A(:, JG, :, 8)
            DENSITY(1:NCL1_io, J, :) = A(:, JG, :, 8)
            Viscousity0(1:NCL1_io, J, :) = A(:, JG, :, 9)
            Viscousity(1:NCL1_io, J, :) = A(:, JG, :, 9)
        END DO
        DO J = 1, N2DO(MYID)
            JG = JDSWT(MYID) - 1 + J
            DO C = 1, 3
                EXPLT0_io(:, J, :, C) = A(:, JG, :, 9 + C)
            END DO
        END DO
        DEALLOCATE(A)
    END SUBROUTINE

    ! SOLVERRK3_MOM_io(NS) (SOLVERRK3_MOM_IO.f90:24-117) for the periodic main domain, either
    ! viscous scheme, calling the reference's routines in the same order; the Poisson
    ! solve is FFT99_POIS3D_periodicxz(IIO) (:92) instead of the compiled default
    ! FISHPACK_POIS3D_SIMPLE(RHSLLPHI_io, DPH_io) (:93), whose output dummy has no y-halo rows
    ! while DPH_io has them, so the default stores its solution shifted in memory.
    SUBROUTINE H_IOSUBSTEP(NS)
        INTEGER(4), INTENT(IN) :: NS
        INTEGER(4) :: IDR
        G0_io(:, :, :, :) = G_io(:, :, :, :)
        CALL VELOCITY_CALC_io
        CALL CONVECTION_X_io
        CALL CONVECTION_Y_io
        CALL CONVECTION_Z_io
        IF (iVisScheme == VisImplicit) THEN
            IDR = 1
            CALL VISCOUS_PAR_EXPLT_X_io
            CALL RHS_MOM_EXPLICIT_io(NS, IDR)
            CALL MOMFA_io(NS, IDR)
            IDR = 2
            CALL VISCOUS_PAR_EXPLT_Y_io
            CALL RHS_MOM_EXPLICIT_io(NS, IDR)
            CALL MOMFA_io(NS, IDR)
            IDR = 3
            CALL VISCOUS_PAR_EXPLT_Z_io
            CALL RHS_MOM_EXPLICIT_io(NS, IDR)
            CALL MOMFA_io(NS, IDR)
        ELSE
            CALL DIVG_U_io
            IDR = 1
            CALL VISCOUS_ALL_EXPLT_X_io
            CALL RHS_MOM_EXPLICIT_io(NS, IDR)
            IDR = 2
            CALL VISCOUS_ALL_EXPLT_Y_io
            CALL RHS_MOM_EXPLICIT_io(NS, IDR)
            IDR = 3
            CALL VISCOUS_ALL_EXPLT_Z_io
            CALL RHS_MOM_EXPLICIT_io(NS, IDR)
            CALL MASSFLUX_CALC_io
        END IF
        CALL INTFC_VARS3(1, NCL1_io, NCL1S, NCL1E, G_io)
        CALL BC_WALL_G_io
        CALL DIVG_io(NS)
        CALL FFT99_POIS3D_periodicxz(IIO)
        CALL INTFC_VARS1(1, NCL1_io, NCL1S, NCL1E, DPH_io)
        CALL BC_WALL_DPH_io
        CALL PRCALC_io(NS)
        CALL MASSFLUX_UPDATE_io(NS)
        CALL VELOCITY_CALC_io
    END SUBROUTINE

    ! Viscousity (current viscosity; the Crank-Nicolson terms read it, the explicit ones read
    ! Viscousity0) from one global L x (M+2) x N file; every rank keeps its y-slab.
    SUBROUTINE H_LOADVISC(fn)
        CHARACTER(*), INTENT(IN) :: fn
        REAL(WP), ALLOCATABLE :: A(:, :, :)
        INTEGER(4) :: J, JG
        ALLOCATE(A(NCL1_io, 0:NCL2 + 1, NCL3))
        OPEN(79, FILE = fn, ACCESS = 'STREAM', FORM = 'UNFORMATTED', STATUS = 'old', &
             ACTION = 'read')
        READ(79) A
        CLOSE(79)
        DO J = 0, N2DO(MYID) + 1
            JG = JDSWT(MYID) - 1 + J
            Viscousity(1:NCL1_io, J, :) = A(:, JG, :)
        END DO
        DEALLOCATE(A)
    END SUBROUTINE

    ! The momentum part of SOLVERRK3_MOM_io(NS) with iVisScheme = 1 (SOLVERRK3_MOM_IO.f90:25,
    ! :27, :38-40, :61-74), stopping before the Poisson step; RHS_io is kept as each MOMFA_io
    ! receives it.
    SUBROUTINE H_IOCN(NS, fn)
        INTEGER(4), INTENT(IN) :: NS
        CHARACTER(*), INTENT(IN) :: fn
        REAL(WP), ALLOCATABLE :: R(:, :, :, :)
        INTEGER(4) :: IDR
        ALLOCATE(R(NCL1_io, N2DO(MYID), NCL3, 3))
        G0_io(:, :, :, :) = G_io(:, :, :, :)
        CALL VELOCITY_CALC_io
        CALL CONVECTION_X_io
        CALL CONVECTION_Y_io
        CALL CONVECTION_Z_io
        DO IDR = 1, 3
            IF (IDR == 1) CALL VISCOUS_PAR_EXPLT_X_io
            IF (IDR == 2) CALL VISCOUS_PAR_EXPLT_Y_io
            IF (IDR == 3) CALL VISCOUS_PAR_EXPLT_Z_io
            CALL RHS_MOM_EXPLICIT_io(NS, IDR)
            R(:, :, :, IDR) = RHS_io(:, 1:N2DO(MYID), :)
            CALL MOMFA_io(NS, IDR)
        END DO
        CALL H_OPENRANK(fn, 80)
        WRITE(80) G_io(1:NCL1_io, 0:N2DO(MYID) + 1, :, :), DRHOI_STG(1:NCL1_io, 0:N2DO(MYID) + 1, :, :)
        WRITE(80) R, EXPLT0_io(:, 1:N2DO(MYID), :, :)
        CLOSE(80)
        DEALLOCATE(R)
    END SUBROUTINE

    SUBROUTINE H_DUMP3H(fn, A)
        CHARACTER(*), INTENT(IN) :: fn
        REAL(WP), INTENT(IN) :: A(NCL1_tg, 0:N2DO(0) + 1, NCL3)
        CALL H_OPENRANK(fn, 80)
        WRITE(80) A(:, 1:N2DO(MYID), :)
        CLOSE(80)
    END SUBROUTINE

    SUBROUTINE H_DUMPTABLE(fn)
        CHARACTER(*), INTENT(IN) :: fn
        OPEN(78, FILE = fn, ACCESS = 'STREAM', FORM = 'UNFORMATTED', STATUS = 'replace', &
             ACTION = 'write')
        WRITE(78) INT(N_LIST, 4), INT(IMIN_DH, 4), INT(IMAX_DH, 4), INT(IMIN_H, 4), INT(IMAX_H, 4)
        WRITE(78) LIST_DH, LIST_H, LIST_T, LIST_D, LIST_M, LIST_K, LIST_CP, LIST_B
        WRITE(78) SplineCoeff_DHH_B, SplineCoeff_DHH_C, SplineCoeff_DHH_D
        WRITE(78) SplineCoeff_HT_B, SplineCoeff_HT_C, SplineCoeff_HT_D
        WRITE(78) SplineCoeff_HD_B, SplineCoeff_HD_C, SplineCoeff_HD_D
        WRITE(78) SplineCoeff_HM_B, SplineCoeff_HM_C, SplineCoeff_HM_D
        WRITE(78) SplineCoeff_HK_B, SplineCoeff_HK_C, SplineCoeff_HK_D
        WRITE(78) SplineCoeff_HCp_B, SplineCoeff_HCp_C, SplineCoeff_HCp_D
        WRITE(78) SplineCoeff_HB_B, SplineCoeff_HB_C, SplineCoeff_HB_D
        WRITE(78) T0, H0, D0, M0, K0, CP0, B0
        CLOSE(78)
    END SUBROUTINE

    SUBROUTINE H_PROPLOOKUP(fin, fout)
        CHARACTER(*), INTENT(IN) :: fin, fout
        INTEGER(4) :: n, i
        REAL(WP), ALLOCATABLE :: dh(:), o(:, :)
        OPEN(79, FILE = fin, ACCESS = 'STREAM', FORM = 'UNFORMATTED', STATUS = 'old', ACTION = 'read')
        READ(79) n
        ALLOCATE(dh(n), o(n, 7))
        READ(79) dh
        CLOSE(79)
        DO i = 1, n
            CALL THERM_PROP_UPDATE_FROM_DH(dh(i), o(i, 1), o(i, 2), o(i, 3), o(i, 4), o(i, 5), &
                                           o(i, 6), o(i, 7))
        END DO
        OPEN(80, FILE = fout, ACCESS = 'STREAM', FORM = 'UNFORMATTED', STATUS = 'replace', &
             ACTION = 'write')
        WRITE(80) o
        CLOSE(80)
        DEALLOCATE(dh, o)
    END SUBROUTINE

    SUBROUTINE H_TDMA(kind, fin, fout)
        INTEGER(4), INTENT(IN) :: kind
        CHARACTER(*), INTENT(IN) :: fin, fout
        INTEGER(4) :: isz, jsz, nl
        REAL(WP), ALLOCATABLE :: A(:, :), B(:, :), C(:, :), R(:, :), BC(:, :)
        OPEN(79, FILE = fin, ACCESS = 'STREAM', FORM = 'UNFORMATTED', STATUS = 'old', ACTION = 'read')
        READ(79) isz, jsz
        ALLOCATE(A(isz, jsz), B(isz, jsz), C(isz, jsz), R(isz, jsz))
        READ(79) A, B, C, R
        nl = isz
        IF (kind == 2) nl = jsz
        ALLOCATE(BC(nl, 2))
        BC = 0.0_WP
        IF (kind == 1 .OR. kind == 2) READ(79) BC
        CLOSE(79)
        SELECT CASE (kind)
        CASE (0)
            CALL TDMAIJI_CYC(A, B, C, R, 1, isz, 1, jsz)
        CASE (1)
            CALL TDMAIJJ_nonCYC(A, B, C, R, BC, 1, jsz, 1, isz)
        CASE (2)
            CALL TDMAIJI_nonCYC(A, B, C, R, BC, 1, isz, 1, jsz)
        CASE (3)
            CALL TDMAIJJ_CYC(A, B, C, R, 1, jsz, 1, isz)
        END SELECT
        OPEN(80, FILE = fout, ACCESS = 'STREAM', FORM = 'UNFORMATTED', STATUS = 'replace', &
             ACTION = 'write')
        WRITE(80) R
        CLOSE(80)
        DEALLOCATE(A, B, C, R, BC)
    END SUBROUTINE

END PROGRAM ref_harness
