/*
 * chapsim_hip.h — C-ABI of libchapsim_hip.so, the MI355X (gfx950) implementation of the
 * CHAPSim 1.0 fractional-step RK3/Crank-Nicolson time step for the periodic plane channel
 * (turbulence-generator domain, flow only).
 *
 * Plain C, plain pointers and sizes.  Every entry point returns 0 on success or a negative
 * CHS_E* code; chs_last_error() gives the message.  The Fortran binding is
 * chapsim1.0_amd/fortran/chapsim_hip_mod.f90 (ISO_C_BINDING).
 *
 * Array layout is the reference's (MEMALLOCT.f90:96-103): Fortran column-major, x fastest,
 * then y, then z, then component, FP64:
 *   Q_tg(NCL1, 0:NCL2+1, NCL3, 3)   velocity u, v, w with one y-ghost row each side
 *   PR_tg(NCL1, 0:NCL2+1, NCL3)     pressure
 *   CONVH0_tg(NCL1, NCL2, NCL3, 3)  previous-substep convection terms (RK3 history)
 * Host arrays are copied; the library keeps no host pointer after a call returns.
 */
#ifndef CHAPSIM_HIP_H
#define CHAPSIM_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    CHS_OK = 0,
    CHS_E_ARG = -1,        /* bad argument or size                                       */
    CHS_E_HIP = -2,        /* HIP runtime error                                          */
    CHS_E_NOMEM = -3,      /* device allocation failed                                   */
    CHS_E_NONFINITE = -4,  /* non-finite value found by chs_check_finite                 */
    CHS_E_STATE = -5,      /* call out of order (e.g. no fields uploaded)                */
    CHS_E_RANGE = -6       /* value outside the property table (the reference STOPs,
                              thermo_properties_table_searching.f90:688-701)               */
};

/* Arithmetic form of the substep. */
enum {
    CHS_PLAIN_FP64 = 0,    /* everything FP64                                             */
    CHS_MIXED = 1          /* FP32 stencils / CN / FFTs, FP64 state, low Poisson modes,
                              reductions and divergence refinement                        */
};

/* Mesh and coefficient arrays as the reference's set-up computes them (CONSPARA,
 * COORDJR, TDMA_COEF, FFT99_ROOT, CONS_RKCOEF).  Index base as in the reference array,
 * counts in brackets:
 *   dyfi  DYFI(1:NCL2)                 [M]     COORDJR.f90:121-123
 *   dyci  DYCI(1:NND2)                 [M+1]   COORDJR.f90:125-145
 *   wff   YCL2ND_WFF(1:NND2)           [M+1]   COORDJR.f90:425-442
 *   wfb   YCL2ND_WFB(1:NND2)           [M+1]
 *   amvr, acvr, apvr  A*VR(1:NND2,1:3) [(M+1)*3] TDMA_COEF.f90:67-104 (column-major)
 *   amph, acph, apph  A*PH(1:NCL2)     [M]     TDMA_COEF.f90:44-65
 *   ak1   AK1(1:NCL1)                  [L]     possion3d_FFT99.f90:186-199
 *   ak3   AK3(1:NCL3)                  [N]     possion3d_FFT99.f90:171-184
 *   tgam, troh, talp  (0:3)            [4]     CONS_RKCOEF.f90:21-51
 * Scalars: dxi = 1/DX, dzi = 1/DZ, dxqi, dzqi (CONSPARA.f90:63-73), volm = VOLM_tg
 * (COORDJR.f90:380-388).  RCCI1/2 and RNDI1/2 are 1 for the channel and not passed. */
typedef struct {
    int L, M, N;
    double dxi, dzi, dxqi, dzqi, volm;
    const double *dyfi, *dyci, *wff, *wfb;
    const double *amvr, *acvr, *apvr;
    const double *amph, *acph, *apph;
    const double *ak1, *ak3;
    const double *tgam, *troh, *talp;
} chs_coeffs;

typedef struct chs_ctx chs_ctx;

/* Library version string. */
const char *chs_version(void);

/* Message for the last failing call on ctx (or a global message if ctx is NULL). */
const char *chs_last_error(const chs_ctx *ctx);

/* Host restatement of the reference's channel set-up for the TG domain (READINI mesh keys
 * NCL1_tg, NCL2, NCL3, HX_tg, HZ, STR2; CONSPARA.f90:44-73, COORDJR.f90:44-145,
 * :380-442, TDMA_COEF.f90:44-104, possion3d_FFT99.f90:171-199, CONS_RKCOEF.f90:21-51).
 * Writes into caller buffers sized as in chs_coeffs (scalars through scal[5] =
 * {dxi, dzi, dxqi, dzqi, volm}); then chs_coeffs can point at them. */
int chs_channel_setup(int L, int M, int N, double HX, double HZ, double STR2, double *scal,
                      double *dyfi, double *dyci, double *wff, double *wfb, double *amvr,
                      double *acvr, double *apvr, double *amph, double *acph, double *apph,
                      double *ak1, double *ak3, double *tgam, double *troh, double *talp);

/* Create a solver context on the current HIP device: copies the coefficients, allocates
 * all device state for one channel and precomputes time-invariant tables (Poisson pivots,
 * FFT twiddles).  Replaces the device-side work of MEM_ALLOCAT + FFT99_POIS3D_INIT
 * (MEMALLOCT.f90:96-103, possion3d_FFT99.f90:68-130, :202-239).  L and N must be even with
 * no prime factors other than 2, 3, 5 (lib/fft99/FFTFAX.f90:56-91); M >= 4. */
int chs_create(const chs_coeffs *c, int mode, chs_ctx **out);
void chs_destroy(chs_ctx *ctx);

/* Copy fields in / out (reference layout above).  convh0 may be NULL (then zeroed on
 * upload, as SOLVE.f90:31, or not returned on download). */
int chs_upload_fields(chs_ctx *ctx, const double *q, const double *pr, const double *convh0);
int chs_download_fields(chs_ctx *ctx, double *q, double *pr, double *convh0);

/* One substep of SOLVERRK3_MOM_tg(NS) (SOLVERRK3_MOM_tg.f90:13-56), NS in {0,1,2,3};
 * NS = 0 is the projection step of CALC_INITIALIZATION_tg (INITIAL_FL_FLD_TG.f90:121).
 * dt = DT and cvisc = CVISC = 1/REN of the current step (SOLVE.f90:167-178).  Enqueued on
 * the context's stream; returns without waiting. */
int chs_substep_tg(chs_ctx *ctx, int ns, double dt, double cvisc);

/* One RK3 step: NS = 1, 2, 3 (SOLVE.f90:114-134, TG branch), ending with CFLMM_tg of the new
 * field (CFL_tg, CFL.f90:13-57) formed on the device for the next step's rule (read it with
 * chs_cfl_tg or chs_get_step_state; no host synchronisation inside the step).  With graphs
 * enabled the step's launches are captured once per (dt, cvisc) into a hipGraph and replayed. */
int chs_step_tg(chs_ctx *ctx, double dt, double cvisc);

/* nsteps steps of SOLVE's time loop with the time step chosen on the device (SOLVE.f90:74-134
 * with BCAST_COMM_STEP's rule, :169-178): per step DT = max(min(DT0, 5 DT, CFLGV / CFLMM),
 * DTMIN) from CFLMM_tg of the current field (CFL.f90:13-57) and the last DT, PhyTIME += DT,
 * the Crank-Nicolson tables rebuilt on the device when DT changes, then the three substeps;
 * cvisc = CVISC.  Nothing returns to the host and a captured step graph is never re-made for a
 * new DT.  The reference stops once DT / DT0 < 0.01 (SOLVE.f90:199-207): the device records
 * that in the step state.  Requires chs_set_step_state first.  Returns without waiting. */
int chs_run_tg(chs_ctx *ctx, int nsteps, double dt0, double cflgv, double dtmin, double cvisc);

/* The device step state of chs_run_tg: set DT (the last step's, or DT0 at the start) and
 * PhyTIME, clearing the stop flag (synchronous); read DT, PhyTIME, CFLMM_tg of the current
 * field (NaN if no step or chs_cfl_tg has formed it) and the stop flag (synchronous). */
int chs_set_step_state(chs_ctx *ctx, double dt, double phytime);
int chs_get_step_state(chs_ctx *ctx, double *dt, double *phytime, double *cflmm, int *stopped);

/* Enable (1) or disable (0, default) hipGraph replay of chs_step_tg. */
int chs_set_graph(chs_ctx *ctx, int enable);

/* FP64-residual refinement sweeps of the pressure correction after each substep (0..4): the
 * divergence of the corrected velocity drives another Poisson solve and correction.  Default
 * 0 per substep and 2 for the NS = 0 projection in the mixed form, 0 / 0 in plain FP64. */
int chs_set_refinement(chs_ctx *ctx, int nref_step, int nref_projection);

/* Launch shapes of the substep kernels (no counterpart in the reference, whose loops have no
 * tiling).  Keys and limits: "conv_by" rows per k_conv_rhs block (1, 2, 4, 8, 16; default 4),
 * "conv_bx" wavefronts along x per k_conv_rhs block (1 or 4, conv_bx * conv_by <= 16; 1),
 * "conv_fence" 1 to keep all of a z-plane's loads ahead of its arithmetic in k_conv_rhs (FP64 0,
 * mixed 1),
 * "conv_kc" z-planes marched per k_conv_rhs thread (1..N; 16), "cnx_nlx" x-lines per CN-x
 * block (0 = 2048/L), "cnz_kz" / "cny_ky" march lengths of the z / y CN stencils (FP64 16 / 48,
 * mixed 8 / 32),
 * "fftx_tk" z-planes per x-transform block (0 = derived; even, divides N, TK/2*L <= 2048),
 * "fftz_nl" z-lines per z-transform block (0 = 2048/N; NL*N <= 2048), "nxcd" number of L2
 * domains (XCDs) the k_conv_rhs tiles are grouped for (1..64; FP64 1 = dispatch order, mixed 8),
 * "fft_reg" 1 (default) to run the register-resident FFTs for line lengths 64 ... 1024, 0 for
 * the LDS Stockham transforms, "cn_xz" 1 to run the x and z Crank-Nicolson solves in one kernel
 * where a block can hold whole x-lines (default 0: two kernels, faster on MI355X at 256^2 x
 * 192), "cnxz_by" x-lines per block of that kernel (1, 2, 4; 1), "conv_fx" rows per block
 * (1 default, 2) of k_conv_rhs applying the x Crank-Nicolson inverse itself where L is 64,
 * 128, 192 or 256 (L * conv_fx <= 512, M divisible by conv_fx) and the inverse's band is
 * short, 0 for the separate x kernel (with conv_fx > 0, conv_bx / conv_by do not apply; the
 * mass-flux correction moves after the z-solve, which changes rounding only), "ytdma_bs"
 * threads per block of the Poisson y-solve (64 default, 128, 256) and "ytdma_rb" its rows per
 * load batch (8, 16 default), "ytdma_c" rows per chunk of the chunked y-Thomas (-1 = chosen by
 * the mesh, 0 = unchunked), "conv_dma" rows per block of the direct-to-LDS convection kernel
 * (4 default, 2; 0 = off), "cny3_ky" rows per chunk of the fused three-component y-solve (48;
 * 0 = M, -1 = one launch per component), "cnzy" 1 (default) to run the z and y Crank-Nicolson
 * solves in one pass where L is a multiple of 64 and N of 16, 2 on every mesh (ragged tiles), 0
 * for the two kernels (the pass needs z and y band half-widths of at most 8), "cnzy_ky" its rows
 * per chunk (48; 0 = M), "cnzy_nxcd" the XCDs its tiles
 * are grouped for (8), "divg_nxcd" and "upd_nxcd" the XCDs the divergence
 * transform's and the correction's tiles are grouped for (8 and 1; 1 = dispatch order), "upd_kz"
 * z-planes each correction thread marches (4) and "upd_by" its rows per block (4), "upd_w2" 1
 * (default) for the correction with two points per lane where L is a multiple of 128 (0: one).  Results do not
 * depend on the shapes except through the summation order of the mass-flux partials
 * (conv_by, conv_kc, conv_fx) and conv_fx's rounding.  A captured step graph is re-made on
 * the next step.  CHS_E_ARG for an unknown key or a value out of range. */
int chs_set_tuning(chs_ctx *ctx, const char *key, int value);

/* ---- Batched tridiagonal solvers ------------------------------------------------------ */
/* The reference's four Thomas solvers on device arrays in its Fortran 2-D layout
 * A, B, C, R(IS:IS+isz-1, JS:JS+jsz-1) (I fastest, no padding), R overwritten by the solution,
 * bit-identical to the reference:
 *   kind 0  TDMAIJI_CYC(A,B,C,R, IS,ISZ, JS,JSZ)          TDMAIJI_CYC.f90:67    cyclic along I
 *   kind 1  TDMAIJJ_nonCYC(A,B,C,R, BCJ, JS,JSZ, IS,ISZ)  TDMAIJJ_nonCYC.f90:67 along J, bc = BCJ(isz, 2)
 *   kind 2  TDMAIJI_nonCYC(A,B,C,R, BCI, IS,ISZ, JS,JSZ)  TDMAIJI_nonCYC.f90:67 along I, bc = BCI(jsz, 2)
 *   kind 3  TDMAIJJ_CYC(A,B,C,R, JS,JSZ, IS,ISZ)          TDMAIJJ_CYC.f90:67    cyclic along J
 * bc: the values before the first and after the last point of each line, (nlines, 2)
 * column-major; ignored (may be NULL) for the cyclic kinds.  Runs on the default stream and
 * returns when done.  CHS_E_ARG for a bad kind or size (a line needs 3 points if cyclic, 2
 * otherwise); chs_tdma_last_error() explains. */
int chs_tdma_lines(int kind, int isz, int jsz, const double *a, const double *b, const double *c, double *r,
                   const double *bc);
const char *chs_tdma_last_error(void);

/* chs_tdma_lines without any allocation (no counterpart in the reference): for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops).  Same
 * arguments, kernel and results (bit-identical); its work buffer (3 * points per line * lines
 * doubles, at most 2^21 = 16 MiB) is a static array each GPU holds for the life of the
 * process, so calls that overlap on one GPU race.  CHS_E_ARG beyond that size. */
int chs_tdma_lines_noalloc(int kind, int isz, int jsz, const double *a, const double *b, const double *c, double *r,
                           const double *bc);

/* Device memory helpers for hosts without their own allocator (e.g. the standalone Poisson
 * workload): plain hipMalloc / hipFree / synchronous hipMemcpy (to_device 1: host->device). */
void *chs_dev_malloc(size_t bytes);
int chs_dev_free(void *p);
int chs_memcpy(void *dst, const void *src, size_t bytes, int to_device);

/* Standalone pressure Poisson solve of FFT99_POIS3D_periodicxz (possion3d_FFT99.f90:
 * 271-505) on the context's mesh.  rhs: RHSLLPHI(1:L,1:M,1:N); phi: DPH(1:L,1:M,1:N)
 * (interior rows).  Host arrays; synchronous. */
int chs_poisson(chs_ctx *ctx, const double *rhs, double *phi);

/* The main-domain pressure solve FISHPACK_POIS3D_SIMPLE(RHS, VAR) (possion3d_fishpack.f90:342-453:
 * FFTPACK real transforms in x and z, per-mode y-Thomas TRID0, scale 1/(L N)) on host cell
 * arrays [N][M][L].  The same discrete operator as chs_poisson, with TRID0's gauge for the
 * (0,0) mode: its last row is pinned, phi_hat_00(J=M) = 0 (possion3d_fishpack.f90:329-331;
 * the reference pins it only when the singular pivot rounds below 1e-14, otherwise its
 * constant is set by rounding).  VAR is the array of the routine's declared shape
 * (NCL1_io, NCL2, NCL3), without y-halo rows. */
int chs_poisson_fishpack(chs_ctx *ctx, const double *rhs, double *phi);

/* Same solve on device arrays (cells, N x M x L as RHSLLPHI) on the context's stream, with
 * nref (0..8) refinement sweeps: r = rhs - L phi in FP64, phi += solve(r).  resmax, if not
 * NULL, receives max |rhs - L phi| after the last sweep (then the call waits).  This is
 * the standalone pressure-Poisson workload and the mixed form's low-precision correction. */
int chs_poisson_device(chs_ctx *ctx, const double *rhs, double *phi, int nref, double *resmax);

/* Per-step diagnostics (CFL_tg, CFL.f90:13-57, and DIVGCK_tg, DIVGCK.f90:15-66), computed
 * on the device from the current fields (the CFL that a step already formed is reused);
 * synchronous (a scalar read back).  CHS_E_NONFINITE if the value is not finite (a NaN in
 * the field propagates through the maximum). */
int chs_cfl_tg(chs_ctx *ctx, double dt, double *cflmm);
int chs_max_div(chs_ctx *ctx, double *maxdiv);

/* Wait for all enqueued work on the context's stream. */
int chs_sync(chs_ctx *ctx);

/* x-z mean statistics of the current field, PP_MEAN_ZX_FLOW_TG (processing/PP_MEAN_ZX.f90:
 * 14-307), computed on the device.  stats (host, [M][122], NULL to skip): per row J the
 * statistics s = 0..3 U1xzL_tg (u, v, w, p), 4..6 UPxzL_tg, 7..12 U2xzL_tg, 13..22 U3xzL_tg
 * (the reference's storage index of the upper-triangle products), 23..31 DVDL1xzL_tg (3 x 3,
 * row major: s = 23 + 3 (i-1) + (j-1) for dU_i/dx_j), 32..40 DVDLPxzL_tg, 41..121
 * DVDL2xzL_tg (9 x 9, row major); scal (host, 2, NULL to skip): U1mean_WORK_tg and
 * U1maxx_WORK_tg (:294-305), the bulk velocity and the largest mean u, which the monitor
 * (processing/PP_SCN_MONITOR.f90:147-253) prints with the wall skin friction.  Synchronous;
 * sums in a fixed order (matches the reference to rounding). */
int chs_mean_zx_tg(chs_ctx *ctx, double *stats, double *scal);

/* Device pointers of the resident state (for zero-copy hosts): q (3 halo fields), pr. */
int chs_device_fields(chs_ctx *ctx, double **q, double **pr);

/* Timing aid: HIP events recorded on the context's stream around each launch of the kernel
 * named `only` (NULL: every kernel; a name of no kernel, e.g. "none": recording off) from the
 * next substep on; chs_profile_get returns the average device time (ms) of one named kernel's
 * launches and their count.  Kernel names: conv_rhs, cn_x, cn_z, cn_y, divg_fftx, fftz_fwd,
 * ytdma, fftz_inv, fftx_inv, update, refine_update. */
int chs_profile_reset(chs_ctx *ctx, const char *only);
int chs_profile_get(chs_ctx *ctx, const char *kernel, double *avg_ms, int *count);

/* ---- Ensembles of independent channels ----------------------------------------------- */

typedef struct chs_ens chs_ens;

/* n channels (each its own mesh, chs_coeffs) sharing one stream.  Each channel is a chs_ctx
 * (chs_ens_channel) for upload/download/diagnostics. */
int chs_ens_create(int n, const chs_coeffs *const *coeffs, int mode, chs_ens **out);
void chs_ens_destroy(chs_ens *e);
int chs_ens_channel(chs_ens *e, int i, chs_ctx **ctx);

/* One RK3 step of every channel (dt[i], cvisc[i]) replayed from a single hipGraph, captured
 * once per set of (dt, cvisc); stage-major order (all channels' NS = 1, then NS = 2, 3).
 * By default each channel's own launches go to parallel capture lanes (a channel's result
 * equals its single-channel step bit for bit).  Batched (chs_ens_set_batched): one launch per
 * kernel kind and stage covers every channel, each block running one channel's tile, with the
 * x-solve in its own kernel and the LDS transforms for every mesh, so a channel equals its
 * single-channel step with the tunings conv_fx 0 and fft_reg 0 bit for bit; channels outside
 * the batched kernels' limits (M > 256, a full z or y inverse, refinement sweeps inside a
 * step) make the step use the lanes. */
int chs_ens_step(chs_ens *e, const double *dt, const double *cvisc);

/* Batched launches (1) or each channel's own launches on capture lanes (0, default); no
 * counterpart in the reference. */
int chs_ens_set_batched(chs_ens *e, int batched);
int chs_ens_sync(chs_ens *e);

/* ---- Thermophysical property tables (heated channel, table mode) ---------------------- */

/* Nondimensional property table and cubic-spline coefficients as the reference's set-up
 * holds them (thermal_info, modules.f90:520-690): LIST_DH, LIST_H, LIST_T, LIST_D, LIST_M,
 * LIST_K, LIST_CP, LIST_B (n each, sorted by T) and SplineCoeff_{DHH,HT,HD,HM,HK,HCp,HB}_
 * {B,C,D}; imin_dh .. imax_h are IMIN_DH, IMAX_DH, IMIN_H, IMAX_H (1-based). */
typedef struct {
    int n, imin_dh, imax_dh, imin_h, imax_h;
    const double *dh, *h, *t, *d, *m, *k, *cp, *b;
    const double *dhh_b, *dhh_c, *dhh_d, *ht_b, *ht_c, *ht_d, *hd_b, *hd_c, *hd_d;
    const double *hm_b, *hm_c, *hm_d, *hk_b, *hk_c, *hk_d, *hcp_b, *hcp_c, *hcp_d;
    const double *hb_b, *hb_c, *hb_d;
} chs_prop_table;

typedef struct chs_prop chs_prop;

/* Host restatement of Property_Table_Nondimensionalization + SPLINE_COEFFICIENTS_FOR_DH/H
 * (thermo_properties_table_searching.f90:1206-1346, :977-1129, CUBIC_SPLINE :842-936).
 * rows: n x 8 row-major dimensional table as the file holds it (P H T D M K CP BETA).
 * Out: list[8*n] = DH,H,T,D,M,K,CP,B; spl[21*n] = DHH,HT,HD,HM,HK,HCp,HB x (B,C,D);
 * idx[4] = IMIN_DH, IMAX_DH, IMIN_H, IMAX_H; ref[7] = T0,H0,D0,M0,K0,CP0,B0. */
int chs_prop_table_setup(int n, const double *rows, double T0, double *list, double *spl, int *idx,
                         double *ref);

/* Copy a table to the current GPU. */
int chs_prop_create(const chs_prop_table *t, chs_prop **out);
void chs_prop_destroy(chs_prop *p);
const char *chs_prop_last_error(const chs_prop *p);

/* THERM_PROP_UPDATE_FROM_DH (THERM_PROP_UPDATE.f90:14-67, table branch) for n values of
 * rho*h: out[7*n] = H | T | D | M | K | Cp | B, bit-identical to the reference (including
 * its D <- H->M, M <- H->K, K <- H->D spline assignment, lines 49-51).  Returns CHS_E_RANGE
 * if a value lies outside the table by more than 1e-8.  Host arrays; synchronous. */
int chs_prop_update_from_dh(chs_prop *p, long n, const double *dh, double *out);

/* ---- Energy equation of the heated channel (table properties) ------------------------- */

/* Thermal wall set-up (BC_WALL.f90:484-643; constants from the reference's set-up):
 *   wall_type[w]    1 = fixed heat flux, 2 = fixed temperature (modules.f90:526-527);
 *                   w = 0 bottom (J = 0), 1 top (J = M+1)
 *   cthecd          CTHECD = 1 / (REN PRT0)
 *   dh_min, dh_max  DHmin, DHmax of the table (the update clamps DH 1e-14 inside them)
 *   wallflux        WALLFLUX(1:NCL1, 1:NCL3, 1:2) nondimensional wall heat flux, [2][N][L]
 *                   (I fastest; index 0 bottom, 1 top); read on iso-flux walls
 *   wall_state[w]   fixed-temperature wall values T, H, D, M, K (T_WAL_GV, H_WAL_GV, D_WAL_GV,
 *                   M_WAL_GV, K_WAL_GV)
 *   top_face        reference J of the YCL2ND_WFF / WFB entry the top iso-flux wall reads:
 *                   BC_WALL.f90:598-601 index them with the rank-local N2DO, so a one-rank
 *                   run reads M and a P-rank run the last rank's N2DO; 0 means M */
typedef struct {
    int wall_type[2];
    double cthecd, dh_min, dh_max;
    const double *wallflux;
    double wall_state[2][5];
    int top_face;
} chs_eng_params;

typedef struct chs_eng chs_eng;

/* Thermal state of one channel on the current GPU: DH, H, T, D, M, K, CP (halo layout
 * (N, M+2, L)), the mass flux G (3 x halo), RHS_ENERGY0 (N, M, L) and DENSITY0, Viscousity0
 * for the momentum substep (DH0, which the reference saves but never reads, is not kept). */
int chs_eng_create(const chs_coeffs *mesh, const chs_prop_table *table, const chs_eng_params *p, chs_eng **out);
void chs_eng_destroy(chs_eng *e);
const char *chs_eng_last_error(const chs_eng *e);

/* Host arrays: dh (N, M+2, L), g[3] x (N, M+2, L), rhs0 (N, M, L) or NULL for zero.
 * init_props != 0 then evaluates THERM_PROP_UPDATE on rows 0..M+1 and BC_WALL_THERMAL, as
 * the reference's start-up does before the first substep. */
int chs_eng_upload(chs_eng *e, const double *dh, const double *g, const double *rhs0, int init_props);

/* Replace G by a device array of 3 halo fields (the momentum context's mass flux). */
int chs_eng_set_mass_flux(chs_eng *e, const double *g_dev);

/* SOLVERRK3_ENG_io(NS) (SOLVERRK3_ENG_IO.f90:11-82) for the periodic channel with table
 * properties: DENSITY0, Viscousity0 <- current; RHS_ENERGY_EXPLICIT
 * (RHS_ENERGY_EXPLICIT.f90:11-155); DH += TGAM(NS) dt RHS + TROH(NS) dt RHS0 with the table
 * limiter (lines 34-48); THERM_PROP_UPDATE on J = 1..M; BC_WALL_THERMAL.  FMA contraction
 * off, reference operation order: matches the reference to rounding (1e-12 relative).
 * CHS_E_RANGE if DH leaves the table (the reference STOPs).  Synchronous. */
int chs_substep_eng(chs_eng *e, int ns, double dt);

/* Launch shape of the energy substep's right-hand side (no counterpart in the reference): key
 * "rhs_lds" 1 (default) for the kernel that stages H, K, T, G_x, G_y per z-plane in LDS, 0 for
 * the one reading them per point (identical results); "rhs_kc" its z-planes per block (1..1024,
 * default 32); "rhs_nxcd" the XCDs its tiles are grouped for (1..64, default 8).  CHS_E_ARG for
 * an unknown key or a value out of range. */
int chs_eng_set_tuning(chs_eng *e, const char *key, int value);

/* Copy fields back (NULL skips one): halo arrays (N, M+2, L), rhs0 (N, M, L). */
int chs_eng_download(chs_eng *e, double *dh, double *h, double *t, double *d, double *m, double *k, double *cp,
                     double *rhs0);

/* ---- Restart files (instantaneous TG field, the reference's format) ------------------- */

/* WRT_INSTANT_VARS_TG (processing/WRT_VARS_TG.f90:12-104): four files
 * <dir>/DNS_perixz_INSTANT_T<phytime as ES15.9>_{U,V,W,P}.D, each int32 NCL1, NCL2, NCL3,
 * ITERG, float64 PhyTIME, REN, DT, then the L x M x N interior values in Fortran order.
 * q [3][N][M+2][L], pr [N][M+2][L] host arrays (e.g. from chs_download_fields). */
int chs_restart_write_tg(const char *dir, int L, int M, int N, int iterg, double phytime, double ren, double dt,
                         const double *q, const double *pr);

/* ReStart_INSTANT_VARS_TG(TFM) (processing/RESTART_tg.f90:11-56, READ_3D_VARS
 * processing/READ_INTP_3D_VARS.f90:20-103) into q, pr (interior rows), then the wall ghost
 * rows FLOWSTART sets after a restart (FLOWSTART.f90:37-40).  iterg, phytime (may be NULL)
 * receive ITERG and PhyTIME.  CHS_E_ARG if a file is missing, short or of another mesh. */
int chs_restart_read_tg(const char *dir, double tfm, int L, int M, int N, int *iterg, double *phytime, double *q,
                        double *pr);

/* WRT_INSTANT_VARS_IO (processing/WRT_VARS_IO.f90:16-227), periodic main domain: the same
 * format, files _U, _V, _W, _P from q (Q_io) and pr (PR_io) and, for a thermal run, _T, _D,
 * _H, _M, _K from t, d, h, m, k (TEMPERATURE, DENSITY, ENTHALPY, Viscousity, THERMCONDT), all
 * [N][M+2][L] host arrays (thermal pointers all NULL for a flow-only run). */
int chs_restart_write_io(const char *dir, int L, int M, int N, int iterg, double phytime, double ren, double dt,
                         const double *q, const double *pr, const double *t, const double *d, const double *h,
                         const double *m, const double *k);

/* ReStart_INSTANT_VARS_io(TFM) (processing/RESTART_io.f90:16-131), file part: _U, _V, _W, _P
 * into flow [3][N][M+2][L] and pr, which the reference takes as G_io (iIniFieldType = 0;
 * note that WRT_INSTANT_VARS_IO writes Q_io into those files) or as Q_io (iIniFieldType = 1),
 * with the wall rows (BC_WALL_G_io, BC_WALL_PR_io); for iIniFieldType = 0 and a thermal run
 * also _T, _D, _H into t, d, h and DH = DENSITY * ENTHALPY into dh (rows 1..M; thermal pointers
 * all NULL otherwise).  The properties, their walls and the staggered fields the reference
 * then rebuilds come from chs_eng_upload(dh, g, NULL, 1) and chs_io_upload / chs_io_props. */
int chs_restart_read_io(const char *dir, double tfm, int L, int M, int N, int ini_field_type, int *iterg,
                        double *phytime, double *flow, double *pr, double *t, double *d, double *h, double *dh);
const char *chs_restart_last_error(void);

/* Interpolated start from a coarser field, INITIAL_INTERP_tg (processing/INTERPOLATION_INITIAL.f90:
 * 189-300) with INITIAL_INTERP_MESH (:51-186) and INTP_3D_VARS (processing/READ_INTP_3D_VARS.f90:
 * 106-165): the given field (uo, vo, wo, po: Loo x Mo x Noo interior values, Fortran order, as
 * the restart files hold them; box HXoo x HZoo; y nodes yndo[Mo+1], INTP_COORD_YND.ini) is
 * repeated to cover the target box and interpolated along x, z, then y by cubic splines with
 * zero end slopes (SPLINE / SPLINT, :341-432) onto the channel mesh (L, M, N, HX, HZ, STR2),
 * u and w at their faces, v at the y nodes, p at the cell centres; u is then divided by the
 * bulk velocity (BULK_VELOCITY_TG, INITIAL_FL_FLD_TG.f90:142-166) and the walls set as
 * FLOWSTART does (FLOWSTART.f90:37-40).  Output halo arrays q [3][N][M+2][L], pr [N][M+2][L].
 * Host only. */
int chs_interp_initial_tg(int Loo, int Mo, int Noo, double HXoo, double HZoo, const double *yndo, const double *uo,
                          const double *vo, const double *wo, const double *po, int L, int M, int N, double HX,
                          double HZ, double STR2, double *q, double *pr);

/* Random start of the turbulence-generator channel, random_FL_FLD_TG (INITIAL_FL_FLD_TG.f90:
 * 16-104): IniField_FLOW_tg (:174-244) fills rows 1..M with perturbations of amplitude VPERG,
 * SVPERG where 1 - |y| < 1/4 (random_type 1 iRandomType real, 2 fixed), or with the sine field
 * of random_type 3 (flag_random_sine, incl. its pressure); the x-z mean of every row is removed
 * (INITIAL_MEAN_TG, processing/INITIAL_MEAN.f90:13-122), u gets the laminar profile Vini
 * (LAMPOISLPROF.f90:29-51), v = 0 on the lower wall, u is divided by the bulk velocity
 * (BULK_VELOCITY_TG) and all of Q multiplied by m_over_d = M_inlet / D_inlet (1 for the
 * isothermal generator).  The perturbations of types 1 and 2 come from a seeded counter-based
 * generator (splitmix64 of the point's key) in place of the compiler's RANDOM_NUMBER; type 2
 * keeps the reference's per-point reseeding with 1973 + I + K + J, so its field depends on
 * I + K + J only.  The result is not divergence-free: CALC_INITIALIZATION_tg
 * (INITIAL_FL_FLD_TG.f90:107-140) follows with the NS = 0 substep, whose convection reads the
 * velocity halo rows as the reference leaves them after its ring exchange (INTFC_VARS3, :93):
 * row 0 = row M, row M+1 = row 1; the pressure halo rows are 0.  Output halo arrays
 * q [3][N][M+2][L], pr [N][M+2][L].  Host only. */
int chs_random_field_tg(int L, int M, int N, double HX, double HZ, double STR2, int random_type, double vperg,
                        double svperg, unsigned long long seed, double m_over_d, double *q, double *pr);

/* ---- Main-domain momentum with variable density and viscosity ------------------------ */

/* Physical constants of the main-domain momentum substep (READINI / Flow_State set-up):
 *   cvisc        CVISC = 1/REN
 *   gravity      |iGravity|: 0 none, else the momentum component (1 x, 2 y, 3 z) that gets the
 *                buoyancy term D_STG * F_A (RHS_MOM_EXPLICIT_io.f90:162-226, channel)
 *   f_a          F_A (Flow_State.f90:96-126)
 *   flow_driven  iFlowDriven: 1 constant mass flux (the x-increment's volume mean is removed),
 *                2 constant skin friction Cf_Given (RHS_MOM_EXPLICIT_io.f90:229-265)
 *   cf_given     Cf_Given
 *   vis_scheme   iVisScheme: 0 explicit viscous terms (VISCOUS_ALL_EXPLT_*_io), 1 Crank-Nicolson
 *                (VISCOUS_PAR_EXPLT_*_io + MOMFA_io line solves; SOLVERRK3_MOM_IO.f90:43-78) */
typedef struct {
    double cvisc;
    int gravity;
    double f_a;
    int flow_driven;
    double cf_given;
    int vis_scheme;
} chs_io_params;

typedef struct chs_io chs_io;

/* Main-domain flow state of one periodic channel on the current GPU: G_io, G0_io, Q_io, D_STG,
 * MU_STG (3 halo fields each, [c][N][M+2][L]), PR_io, DENSITY0, Viscousity0 (halo), EXPLT0_io
 * ([3][N][M][L]), DrhoDtP and the Poisson right-hand side (cells), with a Poisson solver on
 * the same mesh (a chs_ctx, FFT99_POIS3D_periodicxz(IIO): the reference's is_FFT_FFT99 method,
 * SOLVERRK3_MOM_IO.f90:92).  Replaces the device side of MEM_ALLOCAT for the main domain
 * (MEMALLOCT.f90:150-201). */
int chs_io_create(const chs_coeffs *mesh, const chs_io_params *p, chs_io **out);
void chs_io_destroy(chs_io *m);
const char *chs_io_last_error(const chs_io *m);

/* Host arrays: g, g0 [3][N][M+2][L] (g0 NULL: = g), pr, d0 (DENSITY0), m0 (Viscousity0)
 * [N][M+2][L], explt0 [3][N][M][L] or NULL for zero.  Synchronous. */
int chs_io_upload(chs_io *m, const double *g, const double *g0, const double *pr, const double *d0,
                  const double *m0, const double *explt0);

/* DENSITY0 and Viscousity0 from device halo arrays (e.g. an energy context's); with the
 * Crank-Nicolson scheme the current viscosity is set to Viscousity0 as well (as chs_io_upload
 * does). */
int chs_io_set_properties(chs_io *m, const double *d0_dev, const double *m0_dev);

/* Crank-Nicolson contexts: the current viscosity Viscousity (host halo array [N][M+2][L]),
 * which VISCOUS_PAR_EXPLT_*_io and MOMFA_io read (the explicit terms read Viscousity0).  In a
 * heated run SOLVERRK3_ENG_io updates it (SOLVERRK3_ENG_IO.f90:23, THERM_PROP_UPDATE);
 * chs_heated_substep takes it from the energy context.  Synchronous. */
int chs_io_upload_viscosity(chs_io *m, const double *visc);

/* What SOLVE runs between the energy and the momentum substep of stage NS in a heated run
 * (SOLVE.f90:122-124): DENSITY_TIME_DERIVATION(NS) (DIVG.f90:249-317), DENSITY_Staggered
 * (MASSFLUX_CALC_IO.f90:262-323), MU_Staggered (:175-259).  Enqueued; returns at once. */
int chs_io_props(chs_io *m, int ns);

/* SOLVERRK3_MOM_io(NS) (SOLVERRK3_MOM_IO.f90:11-120), iWeightedPre = 0: VELOCITY_CALC_io,
 * CONVECTION_X/Y/Z_io; explicit scheme: DIVG_U_io, VISCOUS_ALL_EXPLT_X/Y/Z_io +
 * RHS_MOM_EXPLICIT_io, MASSFLUX_CALC_io; Crank-Nicolson: per component
 * VISCOUS_PAR_EXPLT_X/Y/Z_io + RHS_MOM_EXPLICIT_io + MOMFA_io (MOMFA_io.f90:18-406; the y-solve
 * in its one-rank form, wall viscosities from the line's own z-plane); then DIVG_io, the
 * Poisson solve, PRCALC_io, MASSFLUX_UPDATE_io, VELOCITY_CALC_io, with the walls
 * (BC_WALL_G/Q/PR/DPH_io).  FMA contraction off, reference operation order: matches the
 * reference to 1e-12 (G, Q) / 1e-10 (pressure).  Enqueued; returns at once. */
int chs_substep_io(chs_io *m, int ns, double dt);

/* The momentum part of chs_substep_io alone (SOLVERRK3_MOM_IO.f90:25-80): leaves the new G
 * before the projection (DIVG_io onwards not run), for hosts that split the substep and for
 * checking the momentum solves.  Enqueued. */
int chs_substep_io_momentum(chs_io *m, int ns, double dt);

/* Launch shape of the main-domain momentum kernel, as chs_set_tuning is for the turbulence
 * generator; no counterpart in the reference.  Key "mom_kz": z-planes each block of the
 * LDS-staged momentum kernel marches (default 32; 0 selects the per-point kernel);
 * "mom_by": its rows per block (4 default, or 8); "mom_nxcd": XCDs its tiles are grouped for
 * (default 8; 1 keeps launch order; also used by the staggered properties); "props_kc": z-planes
 * per block of the LDS-staged staggered-property kernel of chs_io_props (default 16; 0 selects
 * the per-point kernel); "mom_rows": rows per 64 x 4 block of the per-point kernel (4
 * default, 2, 1; the block's other rows of threads take further z-planes); "cn_cols": most
 * x-columns per block of the Crank-Nicolson z and y Jacobi solves (1..64, default 16; rounded
 * down to whole wavefronts, the sequential solves where that leaves none), "cnz_cols" and
 * "cny_cols" the same for one direction (0: cn_cols), "cnz_p" / "cny_p" their line points per
 * thread (8 default, or 4); "cnx_wave": the x-solve one wavefront
 * per line (1, default, for L = 64, 128, 256, 512) or one block per line (0); "cn_fastdiv":
 * the z and y solves' coefficients by Newton reciprocals (1, default) or IEEE division (0),
 * which changes them by an ulp, below the 1e-17 the sweeps resolve.  Otherwise results change
 * only through the summation order of the mass-flux partials.  CHS_E_ARG for an unknown key
 * or value. */
int chs_io_set_tuning(chs_io *m, const char *key, int value);

/* Mixed form of the main-domain momentum substep (mixed = 1): EXPLT0_io (the RK history of
 * the explicit terms) and the momentum increment RHS_io, which the line solves transform in
 * place, are kept in FP32; the terms, the solves and every state and property field stay
 * FP64.  mixed = 0 returns to the plain FP64 form.  The current EXPLT0 carries over (rounded
 * to FP32, or widened back); uploads and downloads keep FP64 arrays.  No counterpart in the
 * reference, which computes in FP64 throughout.  Synchronous; CHS_E_ARG for another value. */
int chs_io_set_mixed(chs_io *m, int mixed);

/* Pressure solver of the main-domain substep, the reference's is_FFT_FFT99 / is_FFT_FISHPACK
 * switch (modules.f90:49-50, SOLVERRK3_MOM_IO.f90:89-94): fishpack = 0 (default) solves with
 * FFT99_POIS3D_periodicxz's gauge of the (0,0) mode, 1 with FISHPACK_POIS3D_SIMPLE's (its last
 * row pinned, possion3d_fishpack.f90:329-331; the solver of chs_poisson_fishpack).  The two
 * differ by a constant in DPH and hence in the pressure.  The reference's own FISHPACK call in
 * the substep passes DPH_io(NCL1_io, 0:N2DO+1, NCL3) to the solver's dummy VAR(NCL1_io, N2DO,
 * NCL3), so its solution lands one row shifted (SOLVERRK3_MOM_IO.f90:89-94); that is not
 * reproduced.  CHS_E_ARG for another value. */
int chs_io_set_poisson(chs_io *m, int fishpack);

/* Copy fields back (NULL skips one): g, q, dstg, mstg [3][N][M+2][L]; pr, dph [N][M+2][L]
 * (dph's ghost rows copy the wall cells, BC_WALL_DPH_io); explt0 [3][N][M][L]; drho, rhsp
 * [N][M][L].  Synchronous. */
int chs_io_download(chs_io *m, double *g, double *q, double *pr, double *dph, double *explt0, double *drho,
                    double *rhsp, double *dstg, double *mstg);
int chs_io_sync(chs_io *m);

/* CFL_io (CFL.f90:59-107) on the current Q_io: max over cells of the convective rate and the
 * viscous rate 2 CVISC (DXQI + DZQI + DYFI^2) Viscousity / DENSITY, times dt (CFLMM_io * dt;
 * dt = 1 gives CFLMM_io itself, what SOLVE's dt rule takes, SOLVE.f90:79-85, :177-178).
 * chs_cfl_io reads the context's DENSITY0 and viscosity (Viscousity for the Crank-Nicolson
 * scheme, else Viscousity0), as a flow-only run holds them; chs_heated_cfl the energy
 * context's current DENSITY and Viscousity, as a heated run calls it at the start of a step.
 * Synchronous. */
int chs_cfl_io(chs_io *m, double dt, double *cfl);

/* DIVGCK_io (DIVGCK.f90:60-175, periodic x): max over cells of |div G + DrhoDtP| for the
 * current G and the DrhoDtP of the last chs_io_props (MAXDIVGV_io).  Synchronous. */
int chs_max_div_io(chs_io *m, double *maxdiv);
int chs_heated_cfl(chs_eng *e, chs_io *m, double dt, double *cfl);

/* One stage of the heated channel as SOLVE runs it (SOLVE.f90:119-133): the energy substep
 * with the momentum context's current G, then its DENSITY0 / Viscousity0 (and, for the
 * Crank-Nicolson scheme, Viscousity) into the momentum context, the staggered properties
 * (DRHOI_STG of DENSITY_IMPLICIT_add is identically zero) and SOLVERRK3_MOM_io(NS). */
int chs_heated_substep(chs_eng *e, chs_io *m, int ns, double dt);

#ifdef __cplusplus
}
#endif
#endif /* CHAPSIM_HIP_H */
