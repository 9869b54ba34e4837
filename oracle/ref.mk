# Builds the unmodified CHAPSim 1.0 Fortran reference from its sources where they lie
# (REF, read-only) into oracle/_ref/ (git-ignored).  Test infrastructure only:
#   oracle/_ref/chapsim      the reference program (used as the CPU baseline and for
#                            whole-run parity), run as `mpirun -np 2` or more
#                            (one rank deadlocks in INTFC_VARS3, INTFC_VARS.f90:57-79)
#   oracle/_ref/ref_harness  oracle/ref_harness.f90 linked against the reference
#                            objects: calls single reference routines on given inputs
#   oracle/_ref/chapsim_gpu  the reference program on one rank with the GPU substep (below)
# Usage: make -f oracle/ref.mk [REF=/root/reference] [-j8]
REF     ?= /root/reference
OUT     := oracle/_ref
OBJDIR  := $(OUT)/obj
FC      := /opt/rocm/bin/amdflang
MPIINC  ?= /opt/conda/include
MPILIB  ?= /opt/conda/lib
# Same precision flags as the reference Makefile (Makefile:34), -O2 as its default build
# (Makefile:81-83), no -march so results do not depend on the host CPU.
FFLAGS  := -O2 -fdefault-real-8 -fdefault-double-8 -I$(MPIINC) -module-dir $(OBJDIR) -I$(OBJDIR)
LDFLAGS := -L$(MPILIB) -Wl,-rpath,$(MPILIB) -lmpifort -lmpi

# Object lists and their order follow the reference Makefile (Makefile:93-216); modules.f90
# and the module-defining files come first so later files find their .mod files.
SRC1 := modules MEMALLOCT possion3d_FFT99 possion3d_fishpack BC_COUTLET_ENEG_RK3 \
        BC_COUTLET_MOM_RK3 BC_WALL BC_INLET CFL CONSPARA CONS_RKCOEF CONVECTION_X \
        CONVECTION_Y CONVECTION_Z COORDJR DIVGCK DIVG Flow_State FLOWSTART functions \
        INDEXL2G INDEXSET INITIAL_FL_FLD_TG INITIAL_FL_THEML_FLD_io INTFC_VARS \
        INTFC_THERMAL_io LAMPOISLPROF TDMA_COEF MASSFLUX_CALC_IO mesh_decomp MOMFA_tg \
        MOMFA_io PRESSURE_CORRECTION READINI RHS_CvLpGpS_tg RHS_ENERGY_EXPLICIT \
        RHS_MOM_EXPLICIT_io SOLVE SOLVERRK3_ENG_IO SOLVERRK3_MOM_IO SOLVERRK3_MOM_tg \
        start_mpi TDMAIJI_CYC TDMAIJJ_CYC TDMAIJI_nonCYC TDMAIJJ_nonCYC THERM_PROP_UPDATE \
        thermo_properties_equations thermo_properties_table_searching TRASP23 \
        VELO_CORRECTION VISCOUS_X_EXPLICIT_io VISCOUS_Y_EXPLICIT_io VISCOUS_Z_EXPLICIT_io \
        VMAV WRTHDL
SRC2 := CHK_conservation INITIAL_MEAN INTERPOLATION_INITIAL READ_INTP_3D_VARS POSTPROESS \
        POSTPROESS_integral_instants PP_SCN_MONITOR PP_SPECOSPEC PP_MEAN_T PP_MEAN_Z \
        PP_MEAN_ZX PP_TEC360_DATA_CHECK PP_QuadrantAnalysis WRT_TEC_AVERAGE_XZperiodic_IO \
        WRT_TEC_AVERAGE_Zperiodic_IO WRT_TEC_AVERAGE_XZperiodic_TG WRT_VARS_TG WRT_VARS_IO \
        RESTART_io RESTART_tg randomgen
SRC3 := CFFT99 CFTFAX CFTRIG FACT FAX FFT99A FFT99B FFT99 FFTFAX FFTRIG VPASSM
SRC5 := dsyev2 dsyevc3 dsyevd3 dsyevh3 dsyevj3 dsyevq3 dsyevv3 dsytrd3 slvsec3 sqrabs \
        zheevc3 zheevd3 zheevh3 zheevj3 zheevq3 zheevv3 zhetrd3 cbabk2 cbal cdiv cg comqr2 \
        comqr corth csroot pythag

OBJS := $(addprefix $(OBJDIR)/,$(addsuffix .o,$(SRC1) $(SRC2) $(SRC3))) $(OBJDIR)/fftpack.o \
        $(addprefix $(OBJDIR)/,$(addsuffix .o,$(SRC5)))

all: $(OUT)/chapsim $(OUT)/ref_harness $(OUT)/chapsim_gpu

$(OBJDIR):
	mkdir -p $(OBJDIR)

# Module files are produced as a side effect, so compile strictly in list order.
$(OBJDIR)/.stamp: | $(OBJDIR)
	set -e; for f in $(SRC1); do $(FC) $(FFLAGS) -c $(REF)/src/$$f.f90 -o $(OBJDIR)/$$f.o; done
	set -e; for f in $(SRC2); do $(FC) $(FFLAGS) -c $(REF)/src/processing/$$f.f90 -o $(OBJDIR)/$$f.o; done
	set -e; for f in $(SRC3); do $(FC) $(FFLAGS) -c $(REF)/lib/fft99/$$f.f90 -o $(OBJDIR)/$$f.o; done
	$(FC) -O2 -fdefault-real-8 -fdefault-double-8 -module-dir $(OBJDIR) -c $(REF)/lib/fishpack4.1/fftpack.f -o $(OBJDIR)/fftpack.o
	set -e; for f in $(SRC5); do $(FC) -O2 -c $(REF)/lib/eigen33/$$f.f -o $(OBJDIR)/$$f.o; done
	$(FC) $(FFLAGS) -c $(REF)/src/DNS_THERMAL.f90 -o $(OBJDIR)/DNS_THERMAL.o
	touch $@

$(OUT)/chapsim: $(OBJDIR)/.stamp
	$(FC) -o $@ $(OBJDIR)/DNS_THERMAL.o $(OBJS) $(LDFLAGS)

$(OUT)/ref_harness: $(OBJDIR)/.stamp oracle/ref_harness.f90
	$(FC) $(FFLAGS) -c oracle/ref_harness.f90 -o $(OBJDIR)/ref_harness.o
	$(FC) -o $@ $(OBJDIR)/ref_harness.o $(OBJS) $(LDFLAGS)

# The reference program run as one process on the GPU: every reference object linked
# unchanged, with the linker's --wrap sending SOLVERRK3_MOM_tg, CFL_tg, DIVGCK_tg,
# POSTPROCESS_TG, CALL_TEC360 and INTFC_VARS3/1 to the single-rank, GPU-resident glue
# chapsim1.0_amd/fortran/chs_tg_single_rank.f90 over libchapsim_hip.so.
PKG     := chapsim1.0_amd
comma   := ,
WRAPS   := solverrk3_mom_tg_ cfl_tg_ divgck_tg_ postprocess_tg_ call_tec360_ intfc_vars3_ intfc_vars1_
$(OUT)/chapsim_gpu: $(OBJDIR)/.stamp $(PKG)/fortran/chapsim_hip_mod.f90 $(PKG)/fortran/chs_tg_single_rank.f90 $(PKG)/libchapsim_hip.so
	$(FC) $(FFLAGS) -c $(PKG)/fortran/chapsim_hip_mod.f90 -o $(OBJDIR)/chapsim_hip_mod.o
	$(FC) $(FFLAGS) -c $(PKG)/fortran/chs_tg_single_rank.f90 -o $(OBJDIR)/chs_tg_single_rank.o
	$(FC) -o $@ $(OBJDIR)/DNS_THERMAL.o $(OBJS) $(OBJDIR)/chapsim_hip_mod.o $(OBJDIR)/chs_tg_single_rank.o \
	    $(addprefix -Wl$(comma)--wrap=,$(WRAPS)) -L$(PKG) -lchapsim_hip -Wl,-rpath,'$$ORIGIN/../../$(PKG)' $(LDFLAGS)

clean:
	rm -rf $(OUT)

.PHONY: all clean
