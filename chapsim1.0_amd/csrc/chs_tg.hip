// libchapsim_hip: MI355X (gfx950) implementation of the CHAPSim 1.0 fractional-step
// RK3 / Crank-Nicolson substep for the periodic plane channel (SOLVERRK3_MOM_tg,
// SOLVERRK3_MOM_tg.f90:13-56) behind the C-ABI in include/chapsim_hip.h.
//
// Device layout (one channel): x fastest, then y, then z (the reference's order,
// MEMALLOCT.f90:96-103).  Halo fields (velocity, pressure, phi) keep the y-ghost rows
// J = 0 and J = M+1; cell fields (RHS, history, divergence) keep J = 1..M.  Spectral data
// after the x- and z-transforms is complex [kz][J][kx], kx = 0..L/2.
//
// Per substep (kernels in launch order):
//   k_conv_rhs      convection (CONVECTION_X/Y/Z_tg) + viscous + pressure gradient + RK
//                   combine (RHS_CvLpGpS_tg) for u, v, w at once; FP64 block partial sums
//                   of the u-RHS for the constant-mass-flux correction
//   k_reduce        deterministic FP64 sum of the partials -> DPGRNS
//   k_cn_x, k_cn_z  Crank-Nicolson x and z solves (MOMFA1_X_tg, MOMFA2_Z_tg) applied as
//                   the exact inverse of the constant-coefficient cyclic operator
//   k_cn_y          y solve (MOMFA3_Y_tg) with the exact banded inverse, Q += dQ, wall BCs
//   k_divg_fftx     divergence (DIVG_tg) + forward real x-FFT of the Poisson RHS
//   k_fftz          forward / inverse complex z-FFT
//   k_ytdma         per-wavenumber y-Thomas (TDMAIJJ_nonCYC in FFT99_POIS3D_periodicxz)
//   k_fftx_inv      inverse real x-FFT -> phi, phi ghost rows (BC_WALL_DPH_tg)
//   k_update        VELOUPDT_tg + PRCALC_tg + wall BCs
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/chapsim_hip.h"
#include "chs_fft.h"
#include "chs_fftr.h"

using chs::cpx;

#define CHS_VERSION "chapsim_hip 0.1 (gfx950)"

namespace {

// Maximum that propagates NaN (fmax drops a NaN operand, so a diverged field would report a
// finite CFL or divergence).
__device__ __forceinline__ double nanmax(double a, double b) { return (a > b || a != a) ? a : b; }

// ---------------------------------------------------------------------------------------
// Device-side geometry passed by value to every kernel.
struct Geo {
    int L, M, N, H;              // H = L/2 + 1
    double dxi, dzi, dxqi, dzqi;
    const double *dyfi;          // [M]     DYFI(J) at J-1
    const double *dyci;          // [M+2]   DYCI(J) at J
    const double *wff, *wfb;     // [M+2]   at J
    const double *amvr, *acvr, *apvr;   // [3][M+2] at [c][J]
};

__device__ __forceinline__ size_t hix(const Geo &g, int i, int J, int k)
{
    return ((size_t)k * (g.M + 2) + J) * g.L + i;
}
__device__ __forceinline__ size_t cix(const Geo &g, int i, int J, int k)
{
    return ((size_t)k * g.M + (J - 1)) * g.L + i;
}

// XCD-aware tile order.  The dispatcher hands consecutive workgroup ids to the XCDs in turn
// (id % nxcd), and each XCD has its own L2; tiles that share halo rows should therefore run
// on the same XCD.  Logical tile l = the id's position within its XCD's share, so XCD x gets
// the contiguous tiles [start_x, start_x + count_x) (x fastest, then y, then z).  A bijection
// for any nxcd >= 1 (1 = identity).
// A block's position in its grid.  Kernel bodies (the *_b functions) read it from here rather
// than from the built-ins, so that a body runs under its own launch (own_blk) or as one item
// of a launch that covers many channels' grids (k_multi).
struct Blk {
    unsigned x, y, z, gx, gy, gz;
};
__device__ __forceinline__ Blk own_blk()
{
    return {blockIdx.x, blockIdx.y, blockIdx.z, gridDim.x, gridDim.y, gridDim.z};
}

__device__ __forceinline__ int xcd_tile(int nxcd, const Blk blk)
{
    const int gx = blk.gx, gy = blk.gy;
    const int nb = gx * gy * blk.gz;
    const int p = blk.x + gx * (blk.y + gy * blk.z);
    if (nxcd <= 1) return p;
    const int xcd = p % nxcd, slot = p / nxcd, per = nb / nxcd, rem = nb % nxcd;
    return (xcd < rem) ? xcd * (per + 1) + slot : rem * (per + 1) + (xcd - rem) * per + slot;
}

// A wave-uniform pointer (the caller guarantees it is the same in every lane) taken into scalar
// registers and made opaque to the optimiser, so that a lane's 32-bit offset is added at the
// access (scalar base + vector offset) instead of being folded into a loop-invariant 64-bit
// vector address that would occupy two registers per stream.
template <typename P> __device__ __forceinline__ __attribute__((address_space(1))) P *uniform(P *p)
{
    // readfirstlane returns int: each half goes through unsigned so that the low word is not
    // sign-extended into the high one
    const unsigned long long v = (unsigned long long)p;
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    unsigned long long u = (unsigned long long)lo | ((unsigned long long)hi << 32);
    asm volatile("" : "+s"(u));
    return (__attribute__((address_space(1))) P *)u;
}

// The cell divergence of DIVG_tg (DIVG.f90:31-46) before its 1 / (alpha DT) scaling, in one
// fixed rounding (no contraction) wherever it is formed: the Poisson RHS and PRCALC_tg's
// RHSLLPHI are the same numbers.
__device__ __forceinline__ double divg_cell(double up, double u, double vp, double v, double wp, double w, double dxi,
                                            double dyf, double dzi)
{
#pragma clang fp contract(off)
    return (up - u) * dxi + (vp - v) * dyf + (wp - w) * dzi;
}

// RK scalars of one substep; the time step and viscosity are read from device memory
// (dv[0] = DT, dv[1] = CVISC), so a step whose DT the device chose needs no host value.
struct Sub {
    double tgam, troh, talp;
    const double *dv;
};

// ---------------------------------------------------------------------------------------
// k_conv_rhs: one thread per column (i, J), marching KC z-planes.  Writes rhs[c] = (gamma H +
// zeta H0 + alpha nu Lap q_c) dt - alpha dt grad_c p (RHS_CvLpGpS_tg.f90:95-160) and H0[c] = H;
// v only for J >= 2 (NII = 2 on the bottom rank).  Also block partial sums of rhs_u / DYFI
// (RHS_CvLpGpS_tg.f90:171-178) in FP64.  Values a column needs again on the next plane (its
// own u, v, w and p of plane k, and the plane k+1 values read for the z-differences) are
// carried in registers, so each plane costs 15 loads of plane k, two of plane k-1, 5 of plane
// k+1 and the 3 history values.
// FX: the block holds whole x-lines (L = 64 BXW) of BY rows and each plane's three RHS rows
// go through LDS, where the circulant x-inverse gx (half-width Dx) of MOMFA1_X_tg is applied
// before the store, so no separate x-solve pass is needed.  The mass-flux correction -DPGRNS,
// which MOMFA_tg receives already subtracted (RHS_CvLpGpS_tg.f90:179-201), is then subtracted
// after the z-solve instead: the cyclic x and z factors map a constant to itself, so only the
// rounding changes.  The partial sums use the unsolved u-row, as the reference's.
template <typename W, int BY, int BXW, bool FENCE, bool FX = false>
__device__ __forceinline__ void k_conv_rhs_b(Geo g, Sub s, const double *__restrict__ q,
                                                      const double *__restrict__ pr,
                                                      W *__restrict__ h0, W *__restrict__ rhs,
                                                      double *__restrict__ partial, int KC, int nxcd,
                                                      const W *__restrict__ gx, int Dx, const Blk blk)
{
    // FX: whole x-lines, BY rows; the block's rows are all inside the mesh (M % BY == 0)
    constexpr int LXB = 64 * BXW * BY;               // one plane's row-block of one component
    __shared__ W lxs[FX ? 2 * 3 * LXB : 1];           // two planes' u, v, w rows (double buffer)
    const int lrow = threadIdx.y * 64 * BXW;
    const int tile = __builtin_amdgcn_readfirstlane(xcd_tile(nxcd, blk));
    const int tx = tile % blk.gx, ty = (tile / blk.gx) % blk.gy, tz = tile / (blk.gx * blk.gy);
    const int i = tx * (64 * BXW) + threadIdx.x;
    // blockDim = (64 BXW, BY): a wavefront is one row J, so J-indexed coefficients are uniform
    const int J = __builtin_amdgcn_readfirstlane(ty * BY + threadIdx.y + 1);
    double mf = 0.0;
    const int kbeg = tz * KC, kend = min(g.N, kbeg + KC);
    if (i < g.L && J <= g.M) {
        const int L = g.L, N = g.N;
        const size_t hs = (size_t)N * (g.M + 2) * L;   // halo field size
        const size_t cs = (size_t)N * g.M * L;         // cell field size
        // Row pointers are wave-uniform (J, k uniform): loads take a scalar base and one of
        // three 32-bit lane offsets (i, i+1, i-1), so no per-load 64-bit address.
        const unsigned ic = i, ip = (i == L - 1) ? 0 : i + 1, im = (i == 0) ? L - 1 : i - 1;
        const double *u = q, *v = q + hs, *w = q + 2 * hs;
        auto row = [&](const double *f, int jj, int kk) { return f + ((size_t)kk * (g.M + 2) + jj) * L; };
        auto crow = [&](int kk) { return ((size_t)kk * g.M + (J - 1)) * L; };   // cell row offset
        // State is FP64.  Sums and differences of state values (and the viscous and
        // pressure-gradient stencils, whose rows cancel) are formed in FP64; the convection
        // products run in the work precision W.
        const double wffJ = g.wff[J], wfbJ = g.wfb[J], wffP = g.wff[J + 1], wfbP = g.wfb[J + 1];
        const double dyf = g.dyfi[J - 1], dyc = g.dyci[J];
        const int o1 = g.M + 2, o2 = 2 * (g.M + 2);
        const double apu = g.apvr[J], acu = g.acvr[J], amu = g.amvr[J];
        const double apv = g.apvr[o1 + J], acv = g.acvr[o1 + J], amv = g.amvr[o1 + J];
        const double apw = g.apvr[o2 + J], acw = g.acvr[o2 + J], amw = g.amvr[o2 + J];
        const double sdt = s.dv[0], scv = s.dv[1];
        const W tgam = (W)s.tgam, troh = (W)s.troh, dt = (W)sdt;
        const double coev = s.talp * scv, coep = s.talp * sdt;
        const W cu1 = (W)(0.25 * g.dxi), cu2 = (W)(dyf * 0.5), cu3 = (W)(0.25 * g.dzi);
        const W cv1 = (W)(g.dxi * 0.5), cv2 = (W)(dyc * 0.25), cv3 = (W)(g.dzi * 0.5);
        // carried: plane k-1 (_m) and plane k (_c)
        const int km0 = (kbeg == 0) ? N - 1 : kbeg - 1;
        double u_m = row(u, J, km0)[ic], v_m = row(v, J, km0)[ic], w_m = row(w, J, km0)[ic];
        double p_m = row(pr, J, km0)[ic];
        double u_c = row(u, J, kbeg)[ic], v_c = row(v, J, kbeg)[ic], w_c = row(w, J, kbeg)[ic];
        double wim_c = row(w, J, kbeg)[im], wjm_c = row(w, J - 1, kbeg)[ic];
        double rsum = 0.0;   // sum of rhs_u over the column's planes (row J: one DYFI)
        for (int k = kbeg; k < kend; ++k) {
            const int kp = (k == N - 1) ? 0 : k + 1, km = (k == 0) ? N - 1 : k - 1;
            // Every load of the plane is issued here, before any store, so a wavefront waits
            // for memory once per plane (v's ghost-row reads at J = 1 are in bounds, unused).
            const double *uk = row(u, J, k), *vk = row(v, J, k), *wk = row(w, J, k), *pk = row(pr, J, k);
            // plane k+1
            const double *ukp = row(u, J, kp), *vkp = row(v, J, kp), *wkp = row(w, J, kp);
            const double u_p = ukp[ic], v_p = vkp[ic], w_p = wkp[ic];
            const double wim_p = wkp[im], wjm_p = row(w, J - 1, kp)[ic];
            // plane k
            const double uip = uk[ip], uim = uk[im], ujp = row(u, J + 1, k)[ic];
            const double ujm = row(u, J - 1, k)[ic];
            const double *vjk = row(v, J + 1, k);
            const double vjp = vjk[ic], vimjp = vjk[im], vim = vk[im];
            const double p0 = pk[ic], pim = pk[im];
            const double wip = wk[ip], wjp = row(w, J + 1, k)[ic];
            const double vip = vk[ip], vjm = row(v, J - 1, k)[ic];
            const double uipjm = row(u, J - 1, k)[ip], pjm = row(pr, J - 1, k)[ic];
            // plane k-1 (not carried: register budget)
            const double uip_m = row(u, J, km)[ip], vjp_m = row(v, J + 1, km)[ic];
            const size_t c0 = crow(k);
            const W *hk = h0 + c0;
            W *rk = rhs + c0, *hkw = h0 + c0;
            const W hu = hk[ic], hv = hk[cs + ic], hw = hk[2 * cs + ic];
            // FENCE: keep the compiler from sinking any of these loads toward its use
            if constexpr (FENCE) asm volatile("" ::: "memory");
            // ---- u (CONVECTION_X.f90:31-59) ----
            {
                const W a1 = (W)(uip + u_c), a2 = (W)(uim + u_c);
                const W H11 = (a1 * a1 - a2 * a2) * cu1;
                const W H12 = ((W)(vjp + vimjp) * (W)(wffP * ujp + wfbP * u_c) -
                               (W)(v_c + vim) * (W)(wffJ * u_c + wfbJ * ujm)) * cu2;
                const W H13 = ((W)(w_p + wim_p) * (W)(u_p + u_c) - (W)(w_c + wim_c) * (W)(u_c + u_m)) * cu3;
                const W conv = -(H11 + H12 + H13);
                const W rhsc = tgam * conv + troh * hu;
                const double lap = (uip - 2.0 * u_c + uim) * g.dxqi + (u_p - 2.0 * u_c + u_m) * g.dzqi +
                                   (ujp * apu + u_c * acu + ujm * amu);
                hkw[ic] = conv;
                const W pgm = (W)((p0 - pim) * g.dxi * coep);
                const W r = (rhsc + (W)(coev * lap)) * dt - pgm;
                if constexpr (FX) lxs[(k & 1) * 3 * LXB + lrow + ic] = r;
                else rk[ic] = r;
                rsum += (double)r;
            }
            // ---- v (CONVECTION_Y.f90:37-84), rows J >= 2 (computed on every row so that its
            // loads stay with the others; stored for J >= 2) ----
            {
                const W H21 = ((W)(wffJ * uip + wfbJ * uipjm) * (W)(vip + v_c) -
                               (W)(wffJ * u_c + wfbJ * ujm) * (W)(v_c + vim)) * cv1;
                const W b1 = (W)(vjp + v_c), b2 = (W)(v_c + vjm);
                const W H22 = (b1 * b1 - b2 * b2) * cv2;
                const W H23 = ((W)(wffJ * w_p + wfbJ * wjm_p) * (W)(v_p + v_c) -
                               (W)(wffJ * w_c + wfbJ * wjm_c) * (W)(v_c + v_m)) * cv3;
                const W conv = -(H21 + H22 + H23);
                const W rhsc = tgam * conv + troh * hv;
                const double lap = (vip - 2.0 * v_c + vim) * g.dxqi + (v_p - 2.0 * v_c + v_m) * g.dzqi +
                                   (vjp * apv + v_c * acv + vjm * amv);
                const W pgm = (W)((p0 - pjm) * dyc * coep);
                if (J >= 2) {
                    hkw[cs + ic] = conv;
                    const W r = (rhsc + (W)(coev * lap)) * dt - pgm;
                    if constexpr (FX) lxs[(k & 1) * 3 * LXB + LXB + lrow + ic] = r;
                    else rk[cs + ic] = r;
                }
            }
            // ---- w (CONVECTION_Z.f90:33-69) ----
            {
                const W H31 = ((W)(uip + uip_m) * (W)(wip + w_c) - (W)(u_c + u_m) * (W)(w_c + wim_c)) * cu1;
                const W H32 = ((W)(vjp + vjp_m) * (W)(wffP * wjp + wfbP * w_c) -
                               (W)(v_c + v_m) * (W)(wffJ * w_c + wfbJ * wjm_c)) * cu2;
                const W c1 = (W)(w_p + w_c), c2 = (W)(w_c + w_m);
                const W H33 = (c1 * c1 - c2 * c2) * cu3;
                const W conv = -(H31 + H32 + H33);
                const W rhsc = tgam * conv + troh * hw;
                const double lap = (wip - 2.0 * w_c + wim_c) * g.dxqi + (w_p - 2.0 * w_c + w_m) * g.dzqi +
                                   (wjp * apw + w_c * acw + wjm_c * amw);
                hkw[2 * cs + ic] = conv;
                const W pgm = (W)((p0 - p_m) * g.dzi * coep);
                const W r = (rhsc + (W)(coev * lap)) * dt - pgm;
                if constexpr (FX) lxs[(k & 1) * 3 * LXB + 2 * LXB + lrow + ic] = r;
                else rk[2 * cs + ic] = r;
            }
            if constexpr (FX) {
                // the whole block runs this loop (every thread holds a point of the row); the
                // next plane writes the other buffer, so one barrier per plane suffices
                __syncthreads();
                const W *lb = lxs + (k & 1) * 3 * LXB + lrow;
                constexpr int LX = 64 * BXW;
#pragma unroll
                for (int cc = 0; cc < 3; ++cc) {
                    if (cc == 1 && J == 1) continue;   // v: rows J >= N2I = 2 (MOMFA_tg.f90:34-35)
                    const W *ln = lb + cc * LXB;
                    W acc = gx[0] * ln[ic];
                    for (int d = 1; d <= Dx; ++d) {
                        int a = (int)ic + d, b = (int)ic - d;
                        if (a >= LX) a -= LX;
                        if (b < 0) b += LX;
                        acc += gx[d] * (ln[a] + ln[b]);
                    }
                    rk[cc * cs + ic] = acc;
                }
            }
            u_m = u_c; v_m = v_c; w_m = w_c; p_m = p0;
            u_c = u_p; v_c = v_p; w_c = w_p; wim_c = wim_p; wjm_c = wjm_p;
        }
        mf = rsum / dyf;
    }
    // block reduction of the mass-flux partial (fixed order -> deterministic); a block of
    // 3 * 2^n threads folds its top third first
    constexpr int NTB = 64 * BXW * BY, P2 = (NTB & (NTB - 1)) ? (NTB / 3) * 2 : NTB;
    static_assert((P2 & (P2 - 1)) == 0, "block size 2^n or 3 * 2^n");
    __shared__ double red[NTB];
    const int t = threadIdx.y * (64 * BXW) + threadIdx.x;
    red[t] = mf;
    __syncthreads();
    if (P2 < NTB) {
        if (t < NTB - P2) red[t] += red[t + P2];
        __syncthreads();
    }
    for (int o = P2 / 2; o > 0; o >>= 1) {
        if (t < o) red[t] += red[t + o];
        __syncthreads();
    }
    if (t == 0) partial[tile] = red[0];
}
template <typename W, int BY, int BXW, bool FENCE, bool FX = false>
__global__ __launch_bounds__(64 * BXW * BY) void k_conv_rhs(Geo g, Sub s, const double *__restrict__ q,
                                                      const double *__restrict__ pr,
                                                      W *__restrict__ h0, W *__restrict__ rhs,
                                                      double *__restrict__ partial, int KC, int nxcd,
                                                      const W *__restrict__ gx = nullptr, int Dx = 0)
{
    k_conv_rhs_b<W, BY, BXW, FENCE, FX>(g, s, q, pr, h0, rhs, partial, KC, nxcd, gx, Dx, own_blk());
}

// k_conv_rhs_lds: the k_conv_rhs FX form (whole x-lines, x-solve applied in LDS, same
// arithmetic) with the state staged through LDS, so that the stencils' neighbours cost LDS
// reads instead of global loads.  Block = L = 64 BXW threads along x by BY rows J0..J0+BY-1,
// marching KC z-planes.  For each plane the rows J0-1 .. J0+BY of u, v, w (and p, which needs
// no row J0+BY) are loaded once, as whole coalesced rows, and kept in LDS for two planes: while
// plane k is computed, the buffers hold planes k and k+1 (the x-, y- and z+1 neighbours); the
// plane k-1 values a point needs are its own from the previous plane, kept in registers.  Plane
// k+2 is loaded into registers one plane ahead and written over plane k's buffer after it has
// been read.  Thread row t loads its own row (whose plane k+1 values are its u_p, v_p, w_p) and
// a share of the two halo rows.
template <typename W, int BY, int BXW>
__global__ __launch_bounds__(64 * BXW * BY) void k_conv_rhs_lds(Geo g, Sub s, const double *__restrict__ q,
                                                          const double *__restrict__ pr, W *__restrict__ h0,
                                                          W *__restrict__ rhs, double *__restrict__ partial,
                                                          int KC, int nxcd, const W *__restrict__ gx, int Dx)
{
    constexpr int L = 64 * BXW, R = BY + 2, NT = L * BY;
    constexpr int NH = 7, HPT = (NH + BY - 1) / BY;   // halo row-loads (4 at J0-1, 3 at J0+BY)
    constexpr int XH = 16, LP = L + 2 * XH;            // RHS rows carry XH wrapped points each side
    __shared__ double st[2][4][R][L];                  // [plane buffer][u v w p][row][x]
    __shared__ W xs[3 * BY * LP];                       // one plane's three RHS row-blocks
    const int tile = __builtin_amdgcn_readfirstlane(xcd_tile(nxcd, own_blk()));
    const int ty = (tile / gridDim.x) % gridDim.y, tz = tile / (gridDim.x * gridDim.y);
    // blockDim.x = L is a multiple of 64, so the row t is wave-uniform: scalar branches and bases
    const int i = threadIdx.x, t = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int J0 = ty * BY + 1;
    const int J = __builtin_amdgcn_readfirstlane(J0 + t);
    const int N = g.N, M = g.M;
    const size_t hs = (size_t)N * (M + 2) * L, cs = (size_t)N * M * L;
    const int kbeg = tz * KC, kend = min(N, kbeg + KC);
    const int ip = (i == L - 1) ? 0 : i + 1, im = (i == 0) ? L - 1 : i - 1;
    auto wrapk = [&](int kk) { return kk >= N ? kk - N : (kk < 0 ? kk + N : kk); };
    // row (jj, kk) of field f (u, v, w, p); f may vary between wavefronts (halo shares), so it
    // selects a base rather than indexing an array of pointers
    auto grow = [&](int f, int jj, int kk) {
        const double *b = (f == 3) ? pr : q + (size_t)f * hs;
        return b + ((size_t)kk * (M + 2) + jj) * L;
    };
    // halo item h (h % BY == t): rows J0-1 (u, v, w, p) and J0+BY (u, v, w)
    auto hrow = [&](int h) { return h < 4 ? 0 : BY + 1; };
    auto hfld = [&](int h) { return h < 4 ? h : h - 4; };
    double own[4], hal[HPT];
    auto load_plane = [&](int kk) {
#pragma unroll
        for (int f = 0; f < 4; ++f) own[f] = grow(f, J, kk)[i];
#pragma unroll
        for (int e = 0; e < HPT; ++e) {
            const int h = e * BY + t;
            if (h < NH) hal[e] = grow(hfld(h), J0 - 1 + hrow(h), kk)[i];
        }
    };
    auto store_plane = [&](int b) {
#pragma unroll
        for (int f = 0; f < 4; ++f) st[b][f][t + 1][i] = own[f];
#pragma unroll
        for (int e = 0; e < HPT; ++e) {
            const int h = e * BY + t;
            if (h < NH) st[b][hfld(h)][hrow(h)][i] = hal[e];
        }
    };
    // RHS value of component cc at x = i, with its copies in the wrapped halo points
    auto put = [&](int cc, W v) {
        W *row = xs + (cc * BY + t) * LP + XH;
        row[i] = v;
        if (i < XH) row[L + i] = v;
        if (i >= L - XH) row[i - L] = v;
    };
    double mf = 0.0;
    {
        const double wffJ = g.wff[J], wfbJ = g.wfb[J], wffP = g.wff[J + 1], wfbP = g.wfb[J + 1];
        const double dyf = g.dyfi[J - 1], dyc = g.dyci[J];
        const int o1 = M + 2, o2 = 2 * (M + 2);
        const double apu = g.apvr[J], acu = g.acvr[J], amu = g.amvr[J];
        const double apv = g.apvr[o1 + J], acv = g.acvr[o1 + J], amv = g.amvr[o1 + J];
        const double apw = g.apvr[o2 + J], acw = g.acvr[o2 + J], amw = g.amvr[o2 + J];
        const double sdt = s.dv[0], scv = s.dv[1];
        const W tgam = (W)s.tgam, troh = (W)s.troh, dt = (W)sdt;
        const double coev = s.talp * scv, coep = s.talp * sdt;
        const W cu1 = (W)(0.25 * g.dxi), cu2 = (W)(dyf * 0.5), cu3 = (W)(0.25 * g.dzi);
        const W cv1 = (W)(g.dxi * 0.5), cv2 = (W)(dyc * 0.25), cv3 = (W)(g.dzi * 0.5);
        // prologue: planes kbeg, kbeg+1 into LDS; plane kbeg-1 values of this point into
        // registers; plane kbeg+2 in flight
        const int km0 = wrapk(kbeg - 1);
        double u_m = grow(0, J, km0)[i], v_m = grow(1, J, km0)[i], w_m = grow(2, J, km0)[i];
        double p_m = grow(3, J, km0)[i], uip_m = grow(0, J, km0)[ip], vjp_m = grow(1, J + 1, km0)[i];
        load_plane(kbeg);
        store_plane(0);
        load_plane(wrapk(kbeg + 1));
        store_plane(1);
        if (kbeg + 2 <= kend) load_plane(wrapk(kbeg + 2));
        __syncthreads();
        const int r = t + 1;
        double rsum = 0.0;
        // one plane; CB (the buffer holding plane k) is a compile-time constant so that every
        // LDS address is a lane offset plus an immediate
        auto plane = [&](int k, auto cbc) {
            constexpr int cb = decltype(cbc)::value, nb = cb ^ 1;
            const size_t c0 = ((size_t)k * M + (J - 1)) * L;
            const W hu = h0[c0 + i], hv = h0[cs + c0 + i], hw = h0[2 * cs + c0 + i];
            const double(*U)[L] = st[cb][0], (*V)[L] = st[cb][1], (*Wf)[L] = st[cb][2], (*Pf)[L] = st[cb][3];
            const double(*Wn)[L] = st[nb][2];
            // values every component reads; the others are read from LDS inside each component's
            // block (the fences keep the compiler from hoisting them, which costs registers)
            const double u_c = U[r][i], v_c = V[r][i], w_c = Wf[r][i], p0 = Pf[r][i];
            const double uip = U[r][ip], vjp = V[r + 1][i], ujm = U[r - 1][i], wim_c = Wf[r][im];
            const double wjm_c = Wf[r - 1][i], w_p = Wn[r][i];
            W cu, cv, cw;
            // ---- u (CONVECTION_X.f90:31-59) ----
            {
                const double uim = U[r][im], ujp = U[r + 1][i], vimjp = V[r + 1][im], vim = V[r][im];
                const double u_p = st[nb][0][r][i], wim_p = Wn[r][im], pim = Pf[r][im];
                const W a1 = (W)(uip + u_c), a2 = (W)(uim + u_c);
                const W H11 = (a1 * a1 - a2 * a2) * cu1;
                const W H12 = ((W)(vjp + vimjp) * (W)(wffP * ujp + wfbP * u_c) -
                               (W)(v_c + vim) * (W)(wffJ * u_c + wfbJ * ujm)) * cu2;
                const W H13 = ((W)(w_p + wim_p) * (W)(u_p + u_c) - (W)(w_c + wim_c) * (W)(u_c + u_m)) * cu3;
                const W conv = -(H11 + H12 + H13);
                const W rhsc = tgam * conv + troh * hu;
                const double lap = (uip - 2.0 * u_c + uim) * g.dxqi + (u_p - 2.0 * u_c + u_m) * g.dzqi +
                                   (ujp * apu + u_c * acu + ujm * amu);
                cu = conv;
                const W pgm = (W)((p0 - pim) * g.dxi * coep);
                const W rr = (rhsc + (W)(coev * lap)) * dt - pgm;
                put(0, rr);
                rsum += (double)rr;
            }
            asm volatile("" ::: "memory");
            // ---- v (CONVECTION_Y.f90:37-84), stored for rows J >= 2 ----
            {
                const double uipjm = U[r - 1][ip], vip = V[r][ip], vim = V[r][im], vjm = V[r - 1][i];
                const double v_p = st[nb][1][r][i], wjm_p = Wn[r - 1][i], pjm = Pf[r - 1][i];
                const W H21 = ((W)(wffJ * uip + wfbJ * uipjm) * (W)(vip + v_c) -
                               (W)(wffJ * u_c + wfbJ * ujm) * (W)(v_c + vim)) * cv1;
                const W b1 = (W)(vjp + v_c), b2 = (W)(v_c + vjm);
                const W H22 = (b1 * b1 - b2 * b2) * cv2;
                const W H23 = ((W)(wffJ * w_p + wfbJ * wjm_p) * (W)(v_p + v_c) -
                               (W)(wffJ * w_c + wfbJ * wjm_c) * (W)(v_c + v_m)) * cv3;
                const W conv = -(H21 + H22 + H23);
                const W rhsc = tgam * conv + troh * hv;
                const double lap = (vip - 2.0 * v_c + vim) * g.dxqi + (v_p - 2.0 * v_c + v_m) * g.dzqi +
                                   (vjp * apv + v_c * acv + vjm * amv);
                cv = conv;
                const W pgm = (W)((p0 - pjm) * dyc * coep);
                put(1, (rhsc + (W)(coev * lap)) * dt - pgm);
            }
            asm volatile("" ::: "memory");
            // ---- w (CONVECTION_Z.f90:33-69) ----
            {
                const double wip = Wf[r][ip], wjp = Wf[r + 1][i];
                const W H31 = ((W)(uip + uip_m) * (W)(wip + w_c) - (W)(u_c + u_m) * (W)(w_c + wim_c)) * cu1;
                const W H32 = ((W)(vjp + vjp_m) * (W)(wffP * wjp + wfbP * w_c) -
                               (W)(v_c + v_m) * (W)(wffJ * w_c + wfbJ * wjm_c)) * cu2;
                const W c1 = (W)(w_p + w_c), c2 = (W)(w_c + w_m);
                const W H33 = (c1 * c1 - c2 * c2) * cu3;
                const W conv = -(H31 + H32 + H33);
                const W rhsc = tgam * conv + troh * hw;
                const double lap = (wip - 2.0 * w_c + wim_c) * g.dxqi + (w_p - 2.0 * w_c + w_m) * g.dzqi +
                                   (wjp * apw + w_c * acw + wjm_c * amw);
                cw = conv;
                const W pgm = (W)((p0 - p_m) * g.dzi * coep);
                put(2, (rhsc + (W)(coev * lap)) * dt - pgm);
            }
            h0[c0 + i] = cu;
            if (J >= 2) h0[cs + c0 + i] = cv;
            h0[2 * cs + c0 + i] = cw;
            u_m = u_c; v_m = v_c; w_m = w_c; p_m = p0; uip_m = uip; vjp_m = vjp;
            __syncthreads();   // RHS rows complete; plane k's buffer read by every thread
            // x-solve (MOMFA1_X_tg): circulant inverse gx of half-width Dx along the line
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) {
                if (cc == 1 && J == 1) continue;   // v: rows J >= N2I = 2 (MOMFA_tg.f90:34-35)
                const W *ln = xs + (cc * BY + t) * LP + XH + i;
                W acc = gx[0] * ln[0];
                for (int d = 1; d <= Dx; ++d) acc += gx[d] * (ln[d] + ln[-d]);   // Dx <= XH
                rhs[cc * cs + c0 + i] = acc;
            }
            if (k + 2 <= kend) {                             // plane kend is the last one read
                store_plane(cb);                             // plane k+2 over plane k
                if (k + 3 <= kend) load_plane(wrapk(k + 3));
            }
            __syncthreads();   // plane k+2 stored; the RHS rows read
        };
        int k = kbeg;
        for (; k + 1 < kend; k += 2) {
            plane(k, std::integral_constant<int, 0>{});
            plane(k + 1, std::integral_constant<int, 1>{});
        }
        if (k < kend) plane(k, std::integral_constant<int, 0>{});
        mf = rsum / dyf;
    }
    // block reduction of the mass-flux partial in a fixed order (xs reused: 3 BY LP W >= NT
    // doubles); a block of 3 * 2^n threads folds its top third first
    constexpr int P2 = (NT & (NT - 1)) ? (NT / 3) * 2 : NT;
    static_assert((P2 & (P2 - 1)) == 0, "block size 2^n or 3 * 2^n");
    double *red = reinterpret_cast<double *>(xs);
    const int tid = t * L + i;
    red[tid] = mf;
    __syncthreads();
    if (P2 < NT) {
        if (tid < NT - P2) red[tid] += red[tid + P2];
        __syncthreads();
    }
    for (int o = P2 / 2; o > 0; o >>= 1) {
        if (tid < o) red[tid] += red[tid + o];
        __syncthreads();
    }
    if (tid == 0) partial[tile] = red[0];
}

// s_waitcnt vmcnt(0), resp. lgkmcnt(0), with the other counters left alone (gfx9 encoding:
// vmcnt in bits 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8)
constexpr int kWaitVm0 = (0x7 << 4) | (0xF << 8);
constexpr int kWaitLgkm0 = 0xF | (0x7 << 4) | (0x3 << 14);

// A workgroup barrier without the memory fence of __syncthreads: the fence would also wait for
// direct-to-LDS loads still in flight into another buffer.  Callers wait for what the barrier
// must cover (lgkmcnt(0) for their LDS stores); the empty asm keeps the compiler from moving
// memory accesses across.
__device__ __forceinline__ void bare_barrier()
{
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// k_conv_rhs_dma: the arithmetic of k_conv_rhs_lds with the state planes moved into LDS by
// direct global-to-LDS loads (global_load_lds_dwordx4: a lane moves 16 bytes and a wavefront
// one 1 KB chunk, with no register in between) into three rotating plane buffers [u v w p]
// [row J0-1 .. J0+BY][x]: while plane k is computed from the buffers of planes k and k+1, plane
// k+2 is on its way into the third (the one plane k-1 used).  With no staging registers a block
// holds BY = 4 rows (1024 threads at L = 256), so the two halo rows add half of the block's own
// row loads where they add seven eighths at BY = 2.  The RHS rows of the x-solve are written
// over plane k's buffer once every wavefront has read it.  Per plane: wait for this
// wavefront's loads of plane k+1 (vmcnt(0)) and barrier (then every wavefront's have landed,
// and plane k-1's buffer is free: its x-solve is done), issue plane k+2, compute, barrier
// (plane k read), write the RHS rows, barrier, x-solve.
template <typename W, int BY, int BXW>
__global__ __launch_bounds__(64 * BXW * BY) void k_conv_rhs_dma(Geo g, Sub s, const double *__restrict__ q,
                                                          const double *__restrict__ pr, W *__restrict__ h0,
                                                          W *__restrict__ rhs, double *__restrict__ partial,
                                                          int KC, int nxcd, const W *__restrict__ gx, int Dx)
{
    constexpr int L = 64 * BXW, R = BY + 2, NT = L * BY, NW = NT / 64;
    constexpr int FS = R * L, PB = 4 * FS;        // doubles per field slab, per plane buffer
    constexpr int NCH = PB / 128;                 // 1 KB chunks per plane buffer
    static_assert(FS % 128 == 0 && NCH % NW == 0, "a chunk lies in one field; chunks split evenly");
    constexpr int CPW = NCH / NW;                 // chunks per wavefront per plane
    constexpr int XH = 16, LP = L + 2 * XH;       // RHS rows carry XH wrapped points each side
    static_assert(3 * BY * LP * sizeof(W) <= PB * sizeof(double) && NT <= PB, "RHS rows fit a buffer");
    __shared__ double b0[PB], b1[PB], b2[PB];
    const int tile = __builtin_amdgcn_readfirstlane(xcd_tile(nxcd, own_blk()));
    const int ty = (tile / gridDim.x) % gridDim.y, tz = tile / (gridDim.x * gridDim.y);
    const int i = threadIdx.x, t = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int lane = i & 63, wv = __builtin_amdgcn_readfirstlane((t * L + i) >> 6);
    const int J0 = ty * BY + 1;
    const int J = __builtin_amdgcn_readfirstlane(J0 + t);
    const int N = g.N, M = g.M;
    const size_t hs = (size_t)N * (M + 2) * L, cs = (size_t)N * M * L;
    const int kbeg = tz * KC, kend = min(N, kbeg + KC);
    const int ip = (i == L - 1) ? 0 : i + 1, im = (i == 0) ? L - 1 : i - 1;
    auto wrapk = [&](int kk) { return kk >= N ? kk - N : (kk < 0 ? kk + N : kk); };
    auto grow = [&](int f, int jj, int kk) {
        const double *b = (f == 3) ? pr : q + (size_t)f * hs;
        return b + ((size_t)kk * (M + 2) + jj) * L;
    };
    // plane kk of rows J0-1 .. J0+BY into buffer B: chunk ch (wave-uniform) is LDS elements
    // ch*128 .. ch*128+127, lane l moves elements ch*128 + 2l and + 2l + 1
    auto dma = [&](double *B, int kk) {
#pragma unroll
        for (int c = 0; c < CPW; ++c) {
            const int ch = wv * CPW + c;
            const int f = (ch * 128) / FS, cf = ch - f * (FS / 128);   // field, chunk within it
            const double *src;
            if constexpr (L % 128 == 0) {   // a chunk is part of one row: uniform base + lane
                constexpr int CR = L / 128;
                src = (const double *)(uniform(grow(f, J0 - 1 + cf / CR, kk) + (cf % CR) * 128) + 2 * lane);
            } else {                        // a chunk spans rows: row and x per lane
                const int e = cf * 128 + 2 * lane, rr = e / L;
                src = grow(f, J0 - 1 + rr, kk) + (e - rr * L);
            }
            __builtin_amdgcn_global_load_lds((const void *)src,
                                             (__attribute__((address_space(3))) void *)(B + ch * 128), 16, 0, 0);
        }
    };
    auto put = [&](W *xs, int cc, W v) {
        W *row = xs + (cc * BY + t) * LP + XH;
        row[i] = v;
        if (i < XH) row[L + i] = v;
        if (i >= L - XH) row[i - L] = v;
    };
    double mf = 0.0;
    {
        const double wffJ = g.wff[J], wfbJ = g.wfb[J], wffP = g.wff[J + 1], wfbP = g.wfb[J + 1];
        const double dyf = g.dyfi[J - 1], dyc = g.dyci[J];
        const int o1 = M + 2, o2 = 2 * (M + 2);
        const double apu = g.apvr[J], acu = g.acvr[J], amu = g.amvr[J];
        const double apv = g.apvr[o1 + J], acv = g.acvr[o1 + J], amv = g.amvr[o1 + J];
        const double apw = g.apvr[o2 + J], acw = g.acvr[o2 + J], amw = g.amvr[o2 + J];
        const double sdt = s.dv[0], scv = s.dv[1];
        const W tgam = (W)s.tgam, troh = (W)s.troh, dt = (W)sdt;
        const double coev = s.talp * scv, coep = s.talp * sdt;
        const W cu1 = (W)(0.25 * g.dxi), cu2 = (W)(dyf * 0.5), cu3 = (W)(0.25 * g.dzi);
        const W cv1 = (W)(g.dxi * 0.5), cv2 = (W)(dyc * 0.25), cv3 = (W)(g.dzi * 0.5);
        // prologue: planes kbeg, kbeg+1 on their way into buffers 0, 1; this point's plane
        // kbeg-1 values into registers
        dma(b0, kbeg);
        dma(b1, wrapk(kbeg + 1));
        const int km0 = wrapk(kbeg - 1);
        double u_m = grow(0, J, km0)[i], v_m = grow(1, J, km0)[i], w_m = grow(2, J, km0)[i];
        double p_m = grow(3, J, km0)[i], uip_m = grow(0, J, km0)[ip], vjp_m = grow(1, J + 1, km0)[i];
        const int r = t + 1;
        double rsum = 0.0;
        W hn[3];
        auto hload = [&](int kk) {
            const size_t cc0 = ((size_t)kk * M + (J - 1)) * L;
#pragma unroll
            for (int f = 0; f < 3; ++f) hn[f] = uniform(h0 + f * cs + cc0)[i];
        };
        hload(kbeg);
        auto plane = [&](int k, auto cbc) {
            constexpr int cb = decltype(cbc)::value;
            double *A = cb == 0 ? b0 : (cb == 1 ? b1 : b2);
            double *Bn = cb == 0 ? b1 : (cb == 1 ? b2 : b0);
            double *Cn = cb == 0 ? b2 : (cb == 1 ? b0 : b1);
            __builtin_amdgcn_s_waitcnt(kWaitVm0);
            __syncthreads();                                  // planes k, k+1 in LDS
            if (k + 2 <= kend) dma(Cn, wrapk(k + 2));         // plane kend is the last one read
            const size_t c0 = ((size_t)k * M + (J - 1)) * L;
            auto hr0 = uniform(h0 + c0), hr1 = uniform(h0 + cs + c0), hr2 = uniform(h0 + 2 * cs + c0);
            const W hu = hn[0], hv = hn[1], hw = hn[2];
            const double *U = A, *V = A + FS, *Wf = A + 2 * FS, *Pf = A + 3 * FS;
            const double *Un = Bn, *Vn = Bn + FS, *Wn = Bn + 2 * FS;
            auto at = [&](const double *F, int rr, int x) { return F[rr * L + x]; };
            const double u_c = at(U, r, i), v_c = at(V, r, i), w_c = at(Wf, r, i), p0 = at(Pf, r, i);
            const double uip = at(U, r, ip), vjp = at(V, r + 1, i), ujm = at(U, r - 1, i), wim_c = at(Wf, r, im);
            const double wjm_c = at(Wf, r - 1, i), w_p = at(Wn, r, i);
            W cu, cv, cw, ru, rv, rw;
            // ---- u (CONVECTION_X.f90:31-59) ----
            {
                const double uim = at(U, r, im), ujp = at(U, r + 1, i), vimjp = at(V, r + 1, im), vim = at(V, r, im);
                const double u_p = at(Un, r, i), wim_p = at(Wn, r, im), pim = at(Pf, r, im);
                const W a1 = (W)(uip + u_c), a2 = (W)(uim + u_c);
                const W H11 = (a1 * a1 - a2 * a2) * cu1;
                const W H12 = ((W)(vjp + vimjp) * (W)(wffP * ujp + wfbP * u_c) -
                               (W)(v_c + vim) * (W)(wffJ * u_c + wfbJ * ujm)) * cu2;
                const W H13 = ((W)(w_p + wim_p) * (W)(u_p + u_c) - (W)(w_c + wim_c) * (W)(u_c + u_m)) * cu3;
                const W conv = -(H11 + H12 + H13);
                const W rhsc = tgam * conv + troh * hu;
                const double lap = (uip - 2.0 * u_c + uim) * g.dxqi + (u_p - 2.0 * u_c + u_m) * g.dzqi +
                                   (ujp * apu + u_c * acu + ujm * amu);
                cu = conv;
                const W pgm = (W)((p0 - pim) * g.dxi * coep);
                ru = (rhsc + (W)(coev * lap)) * dt - pgm;
                rsum += (double)ru;
            }
            asm volatile("" ::: "memory");
            // ---- v (CONVECTION_Y.f90:37-84), stored for rows J >= 2 ----
            {
                const double uipjm = at(U, r - 1, ip), vip = at(V, r, ip), vim = at(V, r, im), vjm = at(V, r - 1, i);
                const double v_p = at(Vn, r, i), wjm_p = at(Wn, r - 1, i), pjm = at(Pf, r - 1, i);
                const W H21 = ((W)(wffJ * uip + wfbJ * uipjm) * (W)(vip + v_c) -
                               (W)(wffJ * u_c + wfbJ * ujm) * (W)(v_c + vim)) * cv1;
                const W b1v = (W)(vjp + v_c), b2v = (W)(v_c + vjm);
                const W H22 = (b1v * b1v - b2v * b2v) * cv2;
                const W H23 = ((W)(wffJ * w_p + wfbJ * wjm_p) * (W)(v_p + v_c) -
                               (W)(wffJ * w_c + wfbJ * wjm_c) * (W)(v_c + v_m)) * cv3;
                const W conv = -(H21 + H22 + H23);
                const W rhsc = tgam * conv + troh * hv;
                const double lap = (vip - 2.0 * v_c + vim) * g.dxqi + (v_p - 2.0 * v_c + v_m) * g.dzqi +
                                   (vjp * apv + v_c * acv + vjm * amv);
                cv = conv;
                const W pgm = (W)((p0 - pjm) * dyc * coep);
                rv = (rhsc + (W)(coev * lap)) * dt - pgm;
            }
            asm volatile("" ::: "memory");
            // ---- w (CONVECTION_Z.f90:33-69) ----
            {
                const double wip = at(Wf, r, ip), wjp = at(Wf, r + 1, i);
                const W H31 = ((W)(uip + uip_m) * (W)(wip + w_c) - (W)(u_c + u_m) * (W)(w_c + wim_c)) * cu1;
                const W H32 = ((W)(vjp + vjp_m) * (W)(wffP * wjp + wfbP * w_c) -
                               (W)(v_c + v_m) * (W)(wffJ * w_c + wfbJ * wjm_c)) * cu2;
                const W c1 = (W)(w_p + w_c), c2 = (W)(w_c + w_m);
                const W H33 = (c1 * c1 - c2 * c2) * cu3;
                const W conv = -(H31 + H32 + H33);
                const W rhsc = tgam * conv + troh * hw;
                const double lap = (wip - 2.0 * w_c + wim_c) * g.dxqi + (w_p - 2.0 * w_c + w_m) * g.dzqi +
                                   (wjp * apw + w_c * acw + wjm_c * amw);
                cw = conv;
                const W pgm = (W)((p0 - p_m) * g.dzi * coep);
                rw = (rhsc + (W)(coev * lap)) * dt - pgm;
            }
            hr0[i] = cu;
            if (J >= 2) hr1[i] = cv;
            hr2[i] = cw;
            u_m = u_c; v_m = v_c; w_m = w_c; p_m = p0; uip_m = uip; vjp_m = vjp;
            bare_barrier();   // plane k read by every wavefront (its reads have returned: their
                              // values are used above), so its buffer takes the RHS rows
            W *xs = reinterpret_cast<W *>(A);
            put(xs, 0, ru);
            put(xs, 1, rv);
            put(xs, 2, rw);
            __builtin_amdgcn_s_waitcnt(kWaitLgkm0);
            bare_barrier();
            // x-solve (MOMFA1_X_tg): circulant inverse gx of half-width Dx along the line
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) {
                if (cc == 1 && J == 1) continue;   // v: rows J >= N2I = 2 (MOMFA_tg.f90:34-35)
                const W *ln = xs + (cc * BY + t) * LP + XH + i;
                W acc = gx[0] * ln[0];
                for (int d = 1; d <= Dx; ++d) acc += gx[d] * (ln[d] + ln[-d]);   // Dx <= XH
                uniform(rhs + cc * cs + c0)[i] = acc;
            }
            // plane k+1's history values, loaded a plane ahead: no vector-memory wait inside a
            // plane is then behind that plane's in-flight state loads
            if (k + 1 < kend) hload(k + 1);
        };
        int k = kbeg;
        for (; k + 2 < kend; k += 3) {
            plane(k, std::integral_constant<int, 0>{});
            plane(k + 1, std::integral_constant<int, 1>{});
            plane(k + 2, std::integral_constant<int, 2>{});
        }
        if (k < kend) plane(k, std::integral_constant<int, 0>{});
        if (k + 1 < kend) plane(k + 1, std::integral_constant<int, 1>{});
        mf = rsum / dyf;
    }
    // block reduction of the mass-flux partial in a fixed order, in buffer 0 once every
    // wavefront is past its last x-solve (no plane load is outstanding: the last plane issues
    // none)
    constexpr int P2 = (NT & (NT - 1)) ? (NT / 3) * 2 : NT;
    static_assert((P2 & (P2 - 1)) == 0, "block size 2^n or 3 * 2^n");
    __syncthreads();
    double *red = b0;
    const int tid = t * L + i;
    red[tid] = mf;
    __syncthreads();
    if (P2 < NT) {
        if (tid < NT - P2) red[tid] += red[tid + P2];
        __syncthreads();
    }
    for (int o = P2 / 2; o > 0; o >>= 1) {
        if (tid < o) red[tid] += red[tid + o];
        __syncthreads();
    }
    if (tid == 0) partial[tile] = red[0];
}

// Crank-Nicolson tables of the RK stages ns (blockIdx.y) for the device's current DT and CVISC
// (dv[0], dv[1]), beta = alpha_ns DT CVISC / 2 (MOMFA_tg.f90:29), computed in FP64 and stored
// in the work precision.
//  block x = 0, 1: first column of the inverse of the periodic x (z) operator 1 - beta dxx,
//    a = c = -beta DXQI (DZQI), b = 1 - a - c (MOMFA_tg.f90:72-82, :129-146).  It is circulant:
//    with r the root of a r^2 + b r + a = 0 inside the unit circle, the entry at cyclic
//    distance d on a line of n points is (r^d + r^(n-d)) / ((1 - r^n) (b + 2 a r)).
//  block x = 2 + comp: band (half-width D) of the inverse of the y operator of the component
//    (MOMFA_tg.f90:206-216): rows J = 1..M, sub a_J = -beta AMVR, diagonal b_J = 1 - beta ACVR,
//    super c_J = -beta APVR, a_1 = c_M = 0 (the zero boundary vectors of TDMAIJJ_nonCYC).  With
//    the forward pivots rho_1 = b_1, rho_J = b_J - a_J c_{J-1} / rho_{J-1} and the backward ones
//    psi_M = b_M, psi_J = b_J - c_J a_{J+1} / psi_{J+1}: inv[J][J] = 1 / (rho_J + psi_J - b_J),
//    inv[I][J] = -(c_I / rho_I) inv[I+1][J] above the diagonal and -(a_I / psi_I) inv[I-1][J]
//    below it (the ratios of Usmani's closed form).  band[J][d + D] = inv[J][J + d], zero
//    outside the matrix.
// stamp[2 ns .. 2 ns + 1] holds the (DT, CVISC) a stage's tables were built for; a stage whose
// stamp equals dv returns at once, so a step graph launches this unconditionally (k_cn_stamp
// records the new values after it).
template <typename W> struct CnDev {
    W *gx[4], *gz[4], *by[4][3];
    int Dy[4][3];
    double talp[4];
};

template <typename W>
__device__ __forceinline__ void k_cn_tables_b(Geo g, CnDev<W> t, const double *__restrict__ dv,
                                                   const double *__restrict__ stamp, int ns0, const Blk blk)
{
    __shared__ double rho[1024], psi[1024];
    const int ns = ns0 + blk.y;
    const double dt = dv[0], cv = dv[1];
    if (stamp[2 * ns] == dt && stamp[2 * ns + 1] == cv) return;
    const double beta = 0.5 * t.talp[ns] * dt * cv;
    const int b = blk.x;
    if (b <= 1) {
        const int n = b == 0 ? g.L : g.N;
        const double a = -beta * (b == 0 ? g.dxqi : g.dzqi), bb = 1.0 - a - a;
        const double r = -2.0 * a / (bb + sqrt(bb * bb - 4.0 * a * a));
        double rn = 1.0;
        for (int m = 0; m < n; ++m) rn *= r;
        const double cc = 1.0 / ((1.0 - rn) * (bb + 2.0 * a * r));
        W *out = b == 0 ? t.gx[ns] : t.gz[ns];
        for (int d = threadIdx.x; d < n; d += blockDim.x) {
            double rd = 1.0, rnd = 1.0;
            for (int m = 0; m < d; ++m) rd *= r;
            for (int m = 0; m < n - d; ++m) rnd *= r;
            out[d] = (W)((rd + rnd) * cc);
        }
        return;
    }
    const int comp = b - 2, M = g.M, D = t.Dy[ns][comp], w = 2 * D + 1;
    W *band = t.by[ns][comp];
    const double *am = g.amvr + comp * (M + 2), *ac = g.acvr + comp * (M + 2), *ap = g.apvr + comp * (M + 2);
    // 0-based row j = J - 1
    auto sub = [&](int j) { return j == 0 ? 0.0 : -beta * am[j + 1]; };
    auto dia = [&](int j) { return 1.0 - beta * ac[j + 1]; };
    auto sup = [&](int j) { return j == M - 1 ? 0.0 : -beta * ap[j + 1]; };
    if (threadIdx.x == 0) {
        rho[0] = dia(0);
        for (int j = 1; j < M; ++j) rho[j] = dia(j) - sub(j) * sup(j - 1) / rho[j - 1];
    } else if (threadIdx.x == 64) {
        psi[M - 1] = dia(M - 1);
        for (int j = M - 2; j >= 0; --j) psi[j] = dia(j) - sup(j) * sub(j + 1) / psi[j + 1];
    }
    __syncthreads();
    for (int col = threadIdx.x; col < M; col += blockDim.x) {
        const double dg = 1.0 / (rho[col] + psi[col] - dia(col));
        band[(size_t)col * w + D] = (W)dg;
        double x = dg;
        for (int i = col - 1; i >= max(0, col - D); --i) {   // rows above: inv[i][col]
            x = -(sup(i) / rho[i]) * x;
            band[(size_t)i * w + (col - i) + D] = (W)x;
        }
        x = dg;
        for (int i = col + 1; i <= min(M - 1, col + D); ++i) {   // rows below
            x = -(sub(i) / psi[i]) * x;
            band[(size_t)i * w + (col - i) + D] = (W)x;
        }
        for (int d = -D; d <= D; ++d)   // entries of row col outside the matrix
            if (col + d < 0 || col + d >= M) band[(size_t)col * w + d + D] = W(0);
    }
}
template <typename W>
__global__ __launch_bounds__(256) void k_cn_tables(Geo g, CnDev<W> t, const double *__restrict__ dv,
                                                   const double *__restrict__ stamp, int ns0)
{
    k_cn_tables_b<W>(g, t, dv, stamp, ns0, own_blk());
}

// Records the (DT, CVISC) the stages ns0 .. ns0 + n - 1 now hold (see k_cn_tables); with
// dv null, marks them as holding nothing (their band layout is about to change).
__device__ __forceinline__ void k_cn_stamp_b(const double *__restrict__ dv, double *__restrict__ stamp, int ns0, int n, const Blk blk)
{
    const int t = threadIdx.x;
    if (t < 2 * n) stamp[2 * ns0 + t] = dv ? dv[t & 1] : __builtin_nan("");
}
__global__ void k_cn_stamp(const double *__restrict__ dv, double *__restrict__ stamp, int ns0, int n)
{
    k_cn_stamp_b(dv, stamp, ns0, n, own_blk());
}

// The time-step rule of BCAST_COMM_STEP (SOLVE.f90:169-178) on the device: from CFLMM of the
// current field (st[2]) and the last DT (st[0]), DT = max(min(DT0, 5 DT, CFLGV / CFLMM), DTMIN);
// CVISC into st[1], PhyTIME += DT (st[3]); st[4] = 1 once DT / DT0 < 0.01, where the reference
// stops (SOLVE.f90:199-207).
__global__ void k_dt_rule(double *__restrict__ st, double dt0, double cflgv, double dtmin, double cvisc)
{
    const int t = threadIdx.x;
    const double dt = fmax(fmin(fmin(dt0, 5.0 * st[0]), cflgv / st[2]), dtmin);
    const double vals[5] = {dt, cvisc, st[2], st[3] + dt, (dt / dt0 < 0.01 || st[4] != 0.0) ? 1.0 : 0.0};
    __syncthreads();
    if (t < 5) st[t] = vals[t];
}

// Host-given DT and CVISC into the device step scalars.
__device__ __forceinline__ void k_set_step_b(double *__restrict__ dv, double dt, double cvisc, const Blk blk)
{
    const int t = threadIdx.x;
    if (t < 2) dv[t] = t == 0 ? dt : cvisc;
}
__global__ void k_set_step(double *__restrict__ dv, double dt, double cvisc)
{
    k_set_step_b(dv, dt, cvisc, own_blk());
}

// Sum of n partials in a fixed order by one block; out[0] = sum * scale.
__device__ __forceinline__ void k_reduce_b(const double *__restrict__ partial, int n, double scale,
                                                double *__restrict__ out, const Blk blk)
{
    __shared__ double red[256];
    double a = 0.0;
    for (int t = threadIdx.x; t < n; t += 256) a += partial[t];
    red[threadIdx.x] = a;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0] * scale;
}
__global__ __launch_bounds__(256) void k_reduce(const double *__restrict__ partial, int n, double scale,
                                                double *__restrict__ out)
{
    k_reduce_b(partial, n, scale, out, own_blk());
}

// ---------------------------------------------------------------------------------------
// Crank-Nicolson line solves as exact inverses.
// x: (1 - beta dxx) on a periodic line is circulant; its inverse is circulant with first
// column gx (computed on the host in long double).  Entries decay geometrically; D is the
// half-width beyond which all entries are below 2^-64 of the diagonal (D = L/2 keeps all).
template <typename W>
__device__ __forceinline__ void k_cn_x_b(Geo g, const W *__restrict__ gx, int D, int full,
                                              W *__restrict__ rhs, const double *__restrict__ dpg,
                                              int NLX, const Blk blk)
{
    // block: NLX consecutive x-lines (J, k) of component blk.y
    extern __shared__ unsigned char smem_raw[];
    W *line = reinterpret_cast<W *>(smem_raw);
    const int c = blk.y;
    const int L = g.L, nl = g.M * g.N;
    const int l0 = blk.x * NLX;
    const size_t cs = (size_t)g.N * g.M * g.L;
    W *r = rhs + c * cs + (size_t)l0 * L;     // line l = k*M + J-1 starts at l*L
    const W sub = (c == 0) ? (W)dpg[0] : W(0);
    const int nel = min(NLX, nl - l0) * L;
    chs::stage_items<8, W>(
        nel, [&](int t) { return r[t]; }, [&](int t, W x) { line[t] = x - sub; });
    __syncthreads();
    for (int t = threadIdx.x; t < nel; t += blockDim.x) {
        const int ln = t / L, i = t - ln * L;
        const int J = (l0 + ln) % g.M + 1;
        if (c == 1 && J == 1) continue;   // v: rows J >= N2I = 2 (MOMFA_tg.f90:34-35)
        const W *lp = line + ln * L;
        W acc;
        if (full) {
            acc = W(0);
            for (int l = 0; l < L; ++l) {
                int d = i - l;
                if (d < 0) d += L;
                acc += gx[d] * lp[l];
            }
        } else {
            acc = gx[0] * lp[i];
            for (int d = 1; d <= D; ++d) {
                int a = i + d, b = i - d;
                if (a >= L) a -= L;
                if (b < 0) b += L;
                acc += gx[d] * (lp[a] + lp[b]);
            }
        }
        r[t] = acc;
    }
}
template <typename W>
__global__ __launch_bounds__(256) void k_cn_x(Geo g, const W *__restrict__ gx, int D, int full,
                                              W *__restrict__ rhs, const double *__restrict__ dpg,
                                              int NLX)
{
    k_cn_x_b<W>(g, gx, D, full, rhs, dpg, NLX, own_blk());
}

// x and z together (MOMFA1_X_tg then MOMFA2_Z_tg): a block holds whole x-lines, one row J per
// 64 LX threads (LX >= L), and marches the z-planes of one chunk of KZ planes in batches of B.
// Each plane's RHS row goes through LDS, where the circulant x-inverse gx is applied; the
// x-solved value enters the thread's window of 2D + 1 planes for the banded z-inverse gz.  The
// chunk's D halo planes on each side are x-solved again (as the separate z-kernel reloads
// them).  u's rows get the mass-flux correction -DPGRNS first; v's row J = 1 is not solved.
// One barrier per batch; LDS holds two batches (double buffer).
template <typename W, int D>
__global__ __launch_bounds__(1024) void k_cn_xz(Geo g, const W *__restrict__ gx, int Dx, int fullx,
                                                const W *__restrict__ gz, const W *__restrict__ rin,
                                                W *__restrict__ rout, const double *__restrict__ dpg, int KZ)
{
    constexpr int B = 8;
    extern __shared__ unsigned char smem_raw[];
    W *buf = reinterpret_cast<W *>(smem_raw);   // [2][B][BY][LX]
    const int L = g.L, N = g.N, LX = blockDim.x, BY = blockDim.y;
    const int i = threadIdx.x, jr = threadIdx.y;
    const int J = __builtin_amdgcn_readfirstlane(blockIdx.x * BY + jr + 1);
    const int c = blockIdx.y % 3, kc = blockIdx.y / 3;
    const bool inrow = J <= g.M && !(c == 1 && J == 1);
    const bool act = inrow && i < L;
    const size_t cs = (size_t)N * g.M * L, ks = (size_t)g.M * L;
    const W *r = rin + c * cs + (inrow ? cix(g, 0, J, 0) : 0);
    W *o = rout + c * cs + (inrow ? cix(g, 0, J, 0) : 0);
    const W sub = (c == 0) ? (W)dpg[0] : W(0);
    W cz[D + 1];
#pragma unroll
    for (int d = 0; d <= D; ++d) cz[d] = gz[d];
    const int k0 = kc * KZ, k1 = min(N, k0 + KZ);
    // planes k0 - D .. k1 + D - 1 (cyclic), x-solved in batches; the first 2D fill the window
    const int np = (k1 - k0) + 2 * D;
    W win[2 * D + 1];
#pragma unroll
    for (int d = 0; d <= 2 * D; ++d) win[d] = W(0);
    for (int p0 = 0, it = 0; p0 < np; p0 += B, ++it) {
        W *bb = buf + (size_t)(it & 1) * B * BY * LX;
        W val[B];
#pragma unroll
        for (int b = 0; b < B; ++b) {
            int kk = k0 - D + p0 + b;
            kk = kk < 0 ? kk + N : (kk >= N ? kk - N : kk);
            val[b] = (act && p0 + b < np) ? r[(size_t)kk * ks + i] - sub : W(0);
        }
#pragma unroll
        for (int b = 0; b < B; ++b) bb[((size_t)b * BY + jr) * LX + i] = val[b];
        __syncthreads();
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const int p = p0 + b;   // position in the plane sequence (uniform)
            if (p >= np) continue;
            const W *ln = bb + ((size_t)b * BY + jr) * LX;
            W xs = W(0);
            if (act) {
                if (fullx) {
                    for (int l = 0; l < L; ++l) {
                        int d = i - l;
                        d = d < 0 ? d + L : d;
                        xs += gx[d] * ln[l];
                    }
                } else {
                    xs = gx[0] * ln[i];
                    for (int d = 1; d <= Dx; ++d) {
                        int a = i + d, e = i - d;
                        a = a >= L ? a - L : a;
                        e = e < 0 ? e + L : e;
                        xs += gx[d] * (ln[a] + ln[e]);
                    }
                }
            }
            // the window is a shift register; once 2D planes are in, its centre is plane
            // k0 + p - 2D
            win[2 * D] = xs;
            if (p >= 2 * D) {
                W acc = cz[0] * win[D];
#pragma unroll
                for (int d = 1; d <= D; ++d) acc += cz[d] * (win[D + d] + win[D - d]);
                if (act) o[(size_t)(k0 + p - 2 * D) * ks + i] = acc;
            }
#pragma unroll
            for (int d = 0; d < 2 * D; ++d) win[d] = win[d + 1];
        }
    }
}

// z: same circulant inverse along k (stride (M)*L in the cell layout), one thread per (i, J).
template <typename W>
__global__ __launch_bounds__(256) void k_cn_z(Geo g, const W *__restrict__ gz, int D, int full,
                                              const W *__restrict__ rin, W *__restrict__ rout,
                                              int comp0)
{
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int J = blockIdx.y * 4 + threadIdx.y + 1;
    const int k = blockIdx.z;
    const int c = comp0;
    if (i >= g.L || J > g.M) return;
    if (c == 1 && J == 1) return;
    const size_t cs = (size_t)g.N * g.M * g.L;
    const W *r = rin + c * cs;
    const int N = g.N;
    W acc = W(0);
    if (full) {
        for (int l = 0; l < N; ++l) {
            int d = k - l;
            if (d < 0) d += N;
            acc += gz[d] * r[cix(g, i, J, l)];
        }
    } else {
        acc = gz[0] * r[cix(g, i, J, k)];
        for (int d = 1; d <= D; ++d) {
            int a = k + d, b = k - d;
            if (a >= N) a -= N;
            if (b < 0) b += N;
            acc += gz[d] * (r[cix(g, i, J, a)] + r[cix(g, i, J, b)]);
        }
    }
    rout[c * cs + cix(g, i, J, k)] = acc;
}

// y: banded exact inverse by[J][d], d = -D..D (row J = 1..M), Q_c += dQ for rows J >= N2I,
// then the wall BCs of BC_WALL_Q_tg (BC_WALL.f90:229-251).  For u the increment of row
// J = 1 is also stored into row 1 of v's RHS, which the reference's shared RHS_tg array
// carries into v's y-solve (RHS_CvLpGpS_tg.f90:45-46 leaves that row, MOMFA3_Y_tg solves it).
template <typename W>
__global__ __launch_bounds__(256) void k_cn_y(Geo g, const W *__restrict__ by, int D,
                                              const W *__restrict__ rin, W *__restrict__ vrow1,
                                              const double *__restrict__ qin, double *__restrict__ qout, int c)
{
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int J = blockIdx.y * 4 + threadIdx.y + 1;
    const int k = blockIdx.z;
    if (i >= g.L || J > g.M) return;
    const size_t cs = (size_t)g.N * g.M * g.L;
    const size_t hs = (size_t)g.N * (g.M + 2) * g.L;
    const W *r = rin + c * cs;
    const double *qi = qin + c * hs;
    double *qc = qout + c * hs;
    const int M = g.M, w = 2 * D + 1;
    if (!(c == 1 && J == 1)) {   // v: rows J >= N2I = 2
        const int lo = J - D < 1 ? 1 : J - D, hi = J + D > M ? M : J + D;
        const W *bj = by + (size_t)(J - 1) * w;
        W acc = W(0);
        for (int l = lo; l <= hi; ++l) acc += bj[l - J + D] * r[cix(g, i, l, k)];
        qc[hix(g, i, J, k)] = qi[hix(g, i, J, k)] + (double)acc;
        if (c == 0 && J == 1 && vrow1) vrow1[cix(g, i, 1, k)] = acc;
    }
    // BC_WALL_Q (bottom and top walls)
    if (J == 1) {
        qc[hix(g, i, 0, k)] = 0.0;
        if (c == 1) qc[hix(g, i, 1, k)] = 0.0;
    }
    if (J == M) qc[hix(g, i, M + 1, k)] = 0.0;
}

// Marching forms of the z and y Crank-Nicolson stencils with the band half-width D a
// compile-time constant (D <= 8), so each input is loaded once and kept in a register window.
//
// z: one thread per (i, J) marches KZ consecutive planes k; circulant inverse gz (cyclic k).
template <typename W, int D>
__device__ __forceinline__ void k_cn_z_march_b(Geo g, const W *__restrict__ gz, const W *__restrict__ rin,
                                                    W *__restrict__ rout, int KZ, const double *__restrict__ dpg, const Blk blk)
{
    const int i = blk.x * 64 + threadIdx.x;
    const int J = __builtin_amdgcn_readfirstlane(blk.y * 4 + threadIdx.y + 1);
    const int c = blk.z % 3, kc = blk.z / 3;
    if (i >= g.L || J > g.M) return;
    if (c == 1 && J == 1) return;
    const int N = g.N;
    const size_t cs = (size_t)N * g.M * g.L;
    const size_t ks = (size_t)g.M * g.L;          // stride between z-planes
    const W *r = rin + c * cs + cix(g, i, J, 0);
    W *o = rout + c * cs + cix(g, i, J, 0);
    // dpg: -DPGRNS of u deferred from an x-solve fused into k_conv_rhs (see there)
    const W sub = (c == 0 && dpg) ? (W)dpg[0] : W(0);
    W coef[D + 1];
#pragma unroll
    for (int d = 0; d <= D; ++d) coef[d] = gz[d];
    const int k0 = kc * KZ, k1 = min(N, k0 + KZ);
    W win[2 * D + 1];
#pragma unroll
    for (int d = -D; d < D; ++d) {
        int kk = k0 + d;
        kk = kk < 0 ? kk + N : (kk >= N ? kk - N : kk);
        win[d + D] = r[(size_t)kk * ks];
    }
    // planes in batches of B: the batch's B loads are in flight together
    constexpr int B = 8;
    auto step = [&](int k, W nx) {
        win[2 * D] = nx;
        W acc = coef[0] * win[D];
#pragma unroll
        for (int d = 1; d <= D; ++d) acc += coef[d] * (win[D + d] + win[D - d]);
        o[(size_t)k * ks] = acc - sub;
#pragma unroll
        for (int d = 0; d < 2 * D; ++d) win[d] = win[d + 1];
    };
    int k = k0;
    for (; k + B <= k1; k += B) {
        W nx[B];
#pragma unroll
        for (int b = 0; b < B; ++b) {
            int kk = k + D + b;
            kk = kk >= N ? kk - N : kk;
            kk = kk >= N ? kk - N : kk;
            nx[b] = r[(size_t)kk * ks];
        }
#pragma unroll
        for (int b = 0; b < B; ++b) step(k + b, nx[b]);
    }
    for (; k < k1; ++k) {
        int kk = k + D;
        if (kk >= N) kk -= N;
        step(k, r[(size_t)kk * ks]);
    }
}
template <typename W, int D>
__global__ __launch_bounds__(256) void k_cn_z_march(Geo g, const W *__restrict__ gz, const W *__restrict__ rin,
                                                    W *__restrict__ rout, int KZ, const double *__restrict__ dpg)
{
    k_cn_z_march_b<W, D>(g, gz, rin, rout, KZ, dpg, own_blk());
}

// y: one thread per (i, k) marches a chunk of KY rows; banded inverse by[J][d] (row J is
// uniform across the wavefront, so its coefficients come through scalar loads); the predicted
// velocity Q + dQ goes to qout (the correction reads it from there and writes Q) with the wall
// BCs (BC_WALL.f90:229-251).  For u the increment of row 1 is stored into row 1 of
// v's input (the reference's shared RHS_tg, RHS_CvLpGpS_tg.f90:45-46).
template <typename W, int D>
__global__ __launch_bounds__(256) void k_cn_y_march(Geo g, const W *__restrict__ by, W *__restrict__ rin,
                                                    const double *__restrict__ qin, double *__restrict__ qout, int c,
                                                    int KY)
{
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int k = blockIdx.y * 4 + threadIdx.y;
    const int jc = blockIdx.z;
    if (i >= g.L || k >= g.N) return;
    const int M = g.M, w = 2 * D + 1;
    const size_t cs = (size_t)g.N * M * g.L, hs = (size_t)g.N * (M + 2) * g.L;
    const W *r = rin + c * cs + cix(g, i, 1, k);   // row J at r[(J-1)*L]
    const double *qi = qin + c * hs;
    double *qc = qout + c * hs;
    const size_t L = g.L;
    const int J0 = (c == 1 && jc == 0) ? 2 : jc * KY + 1;
    const int J1 = min(M, jc * KY + KY);
    W win[2 * D + 1];
#pragma unroll
    for (int d = -D; d < D; ++d) {
        const int jj = J0 + d;
        win[d + D] = (jj >= 1 && jj <= M) ? r[(size_t)(jj - 1) * L] : W(0);
    }
    // rows in batches of B: the batch's input rows and Q values are loaded together
    constexpr int B = 8;
    auto step = [&](int J, W nx, double qv) {
        win[2 * D] = nx;
        const W *bj = by + (size_t)(J - 1) * w;
        W acc = W(0);
#pragma unroll
        for (int d = 0; d < w; ++d) acc += bj[d] * win[d];   // out-of-range rows hold 0 and 0 coefficients
        qc[hix(g, i, J, k)] = qv + (double)acc;
        if (c == 0 && J == 1) rin[cs + cix(g, i, 1, k)] = acc;
#pragma unroll
        for (int d = 0; d < 2 * D; ++d) win[d] = win[d + 1];
    };
    int J = J0;
    for (; J + B - 1 <= J1; J += B) {
        W nx[B];
        double qv[B];
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const int jj = J + b + D;
            nx[b] = jj <= M ? r[(size_t)(jj - 1) * L] : W(0);
            qv[b] = qi[hix(g, i, J + b, k)];
        }
#pragma unroll
        for (int b = 0; b < B; ++b) step(J + b, nx[b], qv[b]);
    }
    for (; J <= J1; ++J) {
        const int jj = J + D;
        step(J, jj <= M ? r[(size_t)(jj - 1) * L] : W(0), qi[hix(g, i, J, k)]);
    }
    if (jc == 0) {
        qc[hix(g, i, 0, k)] = 0.0;
        if (c == 1) qc[hix(g, i, 1, k)] = 0.0;
    }
    if (J1 == M) qc[hix(g, i, M + 1, k)] = 0.0;
}

// The three y-solves in one launch (blockIdx.z = component + 3 chunk), the components' bands
// of one half-width D.  v's row 1 input, u's increment at row 1 in the reference's shared
// RHS_tg (RHS_CvLpGpS_tg.f90:45-46), is formed by the v thread itself from u's input rows, as
// u's thread forms it, so the launch has no order between components.
template <typename W, int D>
__device__ __forceinline__ void k_cn_y3_march_b(Geo g, const W *__restrict__ by0, const W *__restrict__ by1,
                                                     const W *__restrict__ by2, const W *__restrict__ rin,
                                                     const double *__restrict__ qin, double *__restrict__ qout, int KY, const Blk blk)
{
    // a wavefront is 64 consecutive i of one z-plane k: the plane's bases are wave-uniform
    // (scalar registers) and a lane's offset (J - 1) L + i is a 32-bit vector value
    const int i = blk.x * 64 + threadIdx.x;
    const int k = __builtin_amdgcn_readfirstlane(blk.y * 4 + threadIdx.y);
    const int c = blk.z % 3, jc = blk.z / 3;
    if (i >= g.L || k >= g.N) return;
    // the component's band (read-only restrict arguments: its coefficients, uniform per row,
    // come through scalar loads)
    const W *__restrict__ by = c == 0 ? by0 : (c == 1 ? by1 : by2);
    const int M = g.M, w = 2 * D + 1, L = g.L;
    const size_t cs = (size_t)g.N * M * L, hs = (size_t)g.N * (M + 2) * L;
    const auto rb0 = uniform(rin + (size_t)k * M * L), rbc = uniform(rin + c * cs + (size_t)k * M * L);
    const auto qib = uniform(qin + c * hs + (size_t)k * (M + 2) * L);
    const auto qob = uniform(qout + c * hs + (size_t)k * (M + 2) * L);
    auto rrow = [&](int cc, int JJ) { return (cc == 0 ? rb0 : rbc) + (unsigned)((JJ - 1) * L); };
    auto qirow = [&](int JJ) { return qib + (unsigned)(JJ * L); };
    auto qorow = [&](int JJ) { return qob + (unsigned)(JJ * L); };
    const unsigned lane = i;
    const int J0 = (c == 1 && jc == 0) ? 2 : jc * KY + 1;
    const int J1 = min(M, jc * KY + KY);
    W win[2 * D + 1];
#pragma unroll
    for (int d = -D; d < D; ++d) {
        const int jj = J0 + d;
        win[d + D] = (jj >= 1 && jj <= M) ? rrow(c, jj)[lane] : W(0);
    }
    if constexpr (D > 0) {
        if (c == 1 && jc == 0) {
            // u's increment at row 1: rows 1..1+Du of u's input against its band row 1, whose
            // entries below the diagonal are zero
            W a = W(0);
#pragma unroll
            for (int d = 0; d <= D; ++d) a += by0[D + d] * rrow(0, 1 + d)[lane];
            win[D - 1] = a;
        }
    }
    constexpr int B = 8;
    auto step = [&](int J, W nx, double qv) {
        win[2 * D] = nx;
        const W *bj = by + (size_t)(J - 1) * w;
        W acc = W(0);
#pragma unroll
        for (int d = 0; d < w; ++d) acc += bj[d] * win[d];   // out-of-range rows hold 0 and 0 coefficients
        qorow(J)[lane] = qv + (double)acc;
#pragma unroll
        for (int d = 0; d < 2 * D; ++d) win[d] = win[d + 1];
    };
    int J = J0;
    for (; J + B - 1 <= J1; J += B) {
        W nx[B];
        double qv[B];
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const int jj = J + b + D;
            nx[b] = jj <= M ? rrow(c, jj)[lane] : W(0);
            qv[b] = qirow(J + b)[lane];
        }
#pragma unroll
        for (int b = 0; b < B; ++b) step(J + b, nx[b], qv[b]);
    }
    for (; J <= J1; ++J) {
        const int jj = J + D;
        step(J, jj <= M ? rrow(c, jj)[lane] : W(0), qirow(J)[lane]);
    }
    if (jc == 0) {
        qorow(0)[lane] = 0.0;
        if (c == 1) qorow(1)[lane] = 0.0;
    }
    if (J1 == M) qorow(M + 1)[lane] = 0.0;
}
template <typename W, int D>
__global__ __launch_bounds__(256) void k_cn_y3_march(Geo g, const W *__restrict__ by0, const W *__restrict__ by1,
                                                     const W *__restrict__ by2, const W *__restrict__ rin,
                                                     const double *__restrict__ qin, double *__restrict__ qout, int KY)
{
    k_cn_y3_march_b<W, D>(g, by0, by1, by2, rin, qin, qout, KY, own_blk());
}

// The z and y Crank-Nicolson inverses in one pass (MOMFA2_Z_tg, then MOMFA3_Y_tg with
// TDMAIJJ_nonCYC, Q + dQ and BC_WALL_Q), so that the z-solved increments never reach HBM.  A
// block is 64 columns i by KB planes k of one component and one chunk of rows; each thread
// (i, k) marches its chunk with the y-band window in registers, as k_cn_y3_march does, and the
// row entering the window is z-solved on the fly from an LDS slab of that row that holds the
// block's KB planes and Dz planes either side (cyclic in k; the first 2 Dz wavefronts load
// the side planes).  Rows go in batches of BT: the registers of the next batch are loaded
// while the current one is solved, and the batch's slabs go to one of two LDS buffers taken
// in turns, one barrier per batch.  The arithmetic is k_cn_z_march's and k_cn_y3_march's, in
// their order.  v's row 1 input (u's increment at row 1, RHS_CvLpGpS_tg.f90:45-46) is formed
// by the v block of the first chunk from u's rows 1..1+D, z-solved the same way, before its
// own rows.  Slab plane s holds plane (k0 + s - DZM) mod N; Dz <= DZM and 2 Dz <= KB.  With RG
// (ragged last tiles), lanes past L and planes past N fill their slab slots from wrapped planes
// and store nothing; without it L is a multiple of 64 and N of KB (no per-lane conditions).
template <typename W, int D, int KB, int BT, int DZM, bool RG>
__global__ __launch_bounds__(64 * KB, (sizeof(W) == 8 && D >= 7) ? 4 : 8) void k_cn_zy_march(Geo g, const W *__restrict__ gz, int Dz,
                                                         const W *__restrict__ by0, const W *__restrict__ by1,
                                                         const W *__restrict__ by2, const W *__restrict__ rin,
                                                         const double *__restrict__ qin, double *__restrict__ qout,
                                                         int KY, const double *__restrict__ dpg, int nxcd)
{
    constexpr int SP = KB + 2 * DZM, w = 2 * D + 1, BUF = BT * SP * 64;
    __shared__ W S[2 * BUF];
    const int L = g.L, M = g.M, N = g.N;
    const int nkt = (N + KB - 1) / KB, nit = (L + 63) / 64;
    const int tile = __builtin_amdgcn_readfirstlane(xcd_tile(nxcd, own_blk()));
    const int kt = tile % nkt, it = (tile / nkt) % nit, cz = tile / (nkt * nit);
    const int c = cz % 3, jc = cz / 3;
    const int lane = threadIdx.x, kw = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int k0 = kt * KB, k = k0 + kw;
    const bool iv = !RG || it * 64 + lane < L, kv = !RG || k < N;
    const int kown = RG ? k % N : k;
    const size_t cs = (size_t)N * M * L, hs = (size_t)N * (M + 2) * L;
    const W *__restrict__ by = c == 0 ? by0 : (c == 1 ? by1 : by2);
    // the side plane this wavefront loads (first 2 Dz wavefronts) and its slab position
    const bool side = kw < 2 * Dz;
    int kh = kw < Dz ? k0 - Dz + kw : k0 + KB + kw - Dz;
    kh = RG ? ((kh % N) + N) % N : (kh < 0 ? kh + N : (kh >= N ? kh - N : kh));
    const int sh = kw < Dz ? DZM - Dz + kw : DZM + KB + kw - Dz;
    const int so = DZM + kw;
    const W sub0 = dpg ? (W)dpg[0] : W(0);   // u's deferred -DPGRNS (see k_cn_z_march)
    W cz0[DZM + 1];
#pragma unroll
    for (int d = 0; d <= DZM; ++d) cz0[d] = d <= Dz ? gz[d] : W(0);
    const int J0 = (c == 1 && jc == 0) ? 2 : jc * KY + 1;
    const int J1 = min(M, jc * KY + KY);
    // items: P rows of u for v's row-1 input, then the rows entering the window, J0-D .. J1+D
    const int P = (c == 1 && jc == 0) ? D + 1 : 0;
    const int eS = J0 - D, NI = P + (J1 + D - eS + 1);
    // kind 0: zero (outside 1..M), 1: loaded and z-solved, 2: v's row 1 (u's increment)
    auto item = [&](int t, int &cc, int &row, int &kind) {
        if (t < P) { cc = 0; row = 1 + t; kind = 1; return; }
        const int e = eS + (t - P);
        cc = c; row = e;
        kind = (e < 1 || e > M) ? 0 : ((c == 1 && e == 1) ? 2 : 1);
    };
    auto rrow = [&](int cc, int row, int kk) {
        return uniform(rin + cc * cs + ((size_t)kk * M + (row - 1)) * L + it * 64);
    };
    auto qirow = [&](int J) { return uniform(qin + c * hs + ((size_t)k * (M + 2) + J) * L + it * 64); };
    auto qorow = [&](int J) { return uniform(qout + c * hs + ((size_t)k * (M + 2) + J) * L + it * 64); };
    struct Regs { W own[BT], hal[BT]; double qv[BT]; };
    auto fetch = [&](int t0, Regs &r) {
#pragma unroll
        for (int b = 0; b < BT; ++b) {
            const int t = t0 + b;
            if (t >= NI) break;
            int cc, row, kind;
            item(t, cc, row, kind);
            if (kind == 1 && iv) {
                r.own[b] = rrow(cc, row, kown)[lane];
                if (side) r.hal[b] = rrow(cc, row, kh)[lane];
            }
            if (t >= P && iv && kv) {
                const int J = eS + (t - P) - D;
                if (J >= J0) r.qv[b] = qirow(J)[lane];
            }
        }
    };
    W win[2 * D + 1];
    W a = W(0);
    auto batch = [&](int t0, W *Sb, Regs &cur, Regs &nxt) {
#pragma unroll
        for (int b = 0; b < BT; ++b) {
            const int t = t0 + b;
            if (t >= NI) break;
            int cc, row, kind;
            item(t, cc, row, kind);
            if (kind == 1) {
                Sb[(b * SP + so) * 64 + lane] = cur.own[b];
                if (side) Sb[(b * SP + sh) * 64 + lane] = cur.hal[b];
            }
        }
        __syncthreads();
        fetch(t0 + BT, nxt);
#pragma unroll
        for (int b = 0; b < BT; ++b) {
            const int t = t0 + b;
            if (t >= NI) break;
            int cc, row, kind;
            item(t, cc, row, kind);
            W zs = W(0);
            if (kind == 1) {
                const W *sr = Sb + (b * SP + so) * 64 + lane;
                W acc = cz0[0] * sr[0];
#pragma unroll
                for (int d = 1; d <= DZM; ++d)
                    if (d <= Dz) acc += cz0[d] * (sr[d * 64] + sr[-d * 64]);
                zs = acc - (cc == 0 ? sub0 : W(0));
            } else if (kind == 2) {
                zs = a;
            }
            if (t < P) {
                a += by0[D + t] * zs;   // u's band row 1, entries on and above the diagonal
            } else {
                win[2 * D] = zs;
                const int J = row - D;
                if (J >= J0) {
                    const W *bj = by + (size_t)(J - 1) * w;
                    W acc = W(0);
#pragma unroll
                    for (int d = 0; d < w; ++d) acc += bj[d] * win[d];   // out-of-range rows hold 0 and 0 coefficients
                    if (iv && kv) qorow(J)[lane] = cur.qv[b] + (double)acc;
                }
#pragma unroll
                for (int d = 0; d < 2 * D; ++d) win[d] = win[d + 1];
            }
        }
    };
    Regs ra, rb;
    fetch(0, ra);
    for (int t0 = 0; t0 < NI; t0 += 2 * BT) {
        batch(t0, S, ra, rb);
        if (t0 + BT < NI) batch(t0 + BT, S + BUF, rb, ra);
    }
    if (!iv || !kv) return;
    double *q0 = qout + c * hs + (size_t)k * (M + 2) * L + it * 64;
    if (jc == 0) {
        q0[lane] = 0.0;
        if (c == 1) q0[L + lane] = 0.0;
    }
    if (J1 == M) q0[(size_t)(M + 1) * L + lane] = 0.0;
}

// ---------------------------------------------------------------------------------------
// Poisson.  Spectral data is complex, layout [J][kx][kz] (kz fastest), so z-lines are
// contiguous and the per-mode y-Thomas reads consecutive kz across a wavefront.
struct PoisArgs {
    chs::FftPlan px, pz;
    const void *twx, *twz;   // cpx<T>[L], cpx<T>[N]
};

__device__ __forceinline__ size_t six(const Geo &g, int J, int kx, int kz)
{
    return ((size_t)(J - 1) * g.H + kx) * g.N + kz;
}

// k_divg_fftx: block = row J and TK consecutive z-planes k0..k0+TK-1, i.e. TK x-lines taken
// as TK/2 complex lines (planes k, k+1 packed as real and imaginary part).  Divergence
// (DIVG_tg, DIVG.f90:31-46; or the cell field `src` if given) goes to div[] (kept for
// PRCALC_tg) and into LDS; after the forward FFT the two half spectra are separated, scaled
// by 1/(L N) (the 1/L of FFT99 and the 1/N applied after CFFT99, possion3d_FFT99.f90:316,
// :327-334) and written as spec[J][kx][k0..k0+TK-1] (TK consecutive values per kx).
template <typename T>
__device__ __forceinline__ void k_divg_fftx_b(Geo g, PoisArgs pa, const double *__restrict__ dv, double talp,
                                                   const double *__restrict__ q,
                                                   double *__restrict__ div, cpx<T> *__restrict__ spec,
                                                   int TK, const double *__restrict__ src, const Blk blk)
{
    extern __shared__ unsigned char smem_raw[];
    const int L = g.L, H = g.H, ls = L + 1;
    cpx<T> *tw = reinterpret_cast<cpx<T> *>(smem_raw);
    cpx<T> *a = tw + L;
    const cpx<T> *twg = reinterpret_cast<const cpx<T> *>(pa.twx);
    for (int t = threadIdx.x; t < L; t += blockDim.x) tw[t] = twg[t];
    const int J = blk.x + 1, k0 = blk.y * TK;
    const int np = TK / 2;
    const size_t hs = (size_t)g.N * (g.M + 2) * L;
    const double *u = q, *v = q + hs, *w = q + 2 * hs;
    struct D2 { double d[2]; };
    const double dyf = g.dyfi[J - 1];
    const double coe0 = dv ? 1.0 / (dv[0] * talp) : 1.0;   // 1 / (alpha DT) (DIVG_tg)
    chs::stage_items<4, D2>(
        np * L,
        [&](int t) {
            const int p = t / L, i = t - p * L;
            D2 r;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int k = k0 + 2 * p + h;
                if (src) {
                    r.d[h] = src[cix(g, i, J, k)];
                } else {
                    const int ip = (i == L - 1) ? 0 : i + 1, kp = (k == g.N - 1) ? 0 : k + 1;
                    const double d = divg_cell(u[hix(g, ip, J, k)], u[hix(g, i, J, k)], v[hix(g, i, J + 1, k)],
                                               v[hix(g, i, J, k)], w[hix(g, i, J, kp)], w[hix(g, i, J, k)], g.dxi,
                                               dyf, g.dzi);
                    r.d[h] = d * coe0;
                }
            }
            return r;
        },
        [&](int t, const D2 &r) {
            const int p = t / L, i = t - p * L;
            if (!src && div) {
                div[cix(g, i, J, k0 + 2 * p)] = r.d[0];
                div[cix(g, i, J, k0 + 2 * p + 1)] = r.d[1];
            }
            a[p * ls + i] = {(T)r.d[0], (T)r.d[1]};
        });
    __syncthreads();
    chs::fft_lds_inplace<T>(a, ls, np, pa.px, tw, -1, threadIdx.x, blockDim.x);
    const T scale = T(1.0 / ((double)L * (double)g.N));
    for (int t = threadIdx.x; t < H * TK; t += blockDim.x) {
        const int kx = t / TK, ln = t - kx * TK;
        const int p = ln >> 1;
        const cpx<T> Z = a[p * ls + kx];
        const cpx<T> Zc = a[p * ls + ((L - kx) % L)];
        // A = (Z + conj Zc)/2, B = (Z - conj Zc)/(2i)
        cpx<T> o;
        if ((ln & 1) == 0) o = {(Z.x + Zc.x) * T(0.5) * scale, (Z.y - Zc.y) * T(0.5) * scale};
        else o = {(Z.y + Zc.y) * T(0.5) * scale, (Zc.x - Z.x) * T(0.5) * scale};
        spec[six(g, J, kx, k0 + ln)] = o;
    }
}
template <typename T>
__global__ __launch_bounds__(256, 4) void k_divg_fftx(Geo g, PoisArgs pa, const double *__restrict__ dv, double talp,
                                                   const double *__restrict__ q,
                                                   double *__restrict__ div, cpx<T> *__restrict__ spec,
                                                   int TK, const double *__restrict__ src)
{
    k_divg_fftx_b<T>(g, pa, dv, talp, q, div, spec, TK, src, own_blk());
}

// Complex z-transform of NL contiguous lines [J][kx][:] (both directions).
template <typename T>
__device__ __forceinline__ void k_fftz_b(Geo g, PoisArgs pa, cpx<T> *__restrict__ spec, int NL, int sign, const Blk blk)
{
    extern __shared__ unsigned char smem_raw[];
    const int N = g.N, ls = N + 1;
    cpx<T> *tw = reinterpret_cast<cpx<T> *>(smem_raw);
    cpx<T> *a = tw + N;
    const cpx<T> *twg = reinterpret_cast<const cpx<T> *>(pa.twz);
    for (int t = threadIdx.x; t < N; t += blockDim.x) tw[t] = twg[t];
    const int nlines = g.M * g.H;
    const int l0 = blk.x * NL;
    const int nl = min(NL, nlines - l0);
    cpx<T> *base = spec + (size_t)l0 * N;
    chs::stage_items<4, cpx<T>>(
        nl * N, [&](int t) { return base[t]; },
        [&](int t, const cpx<T> &x) {
            const int ln = t / N, kz = t - ln * N;
            a[ln * ls + kz] = x;
        });
    __syncthreads();
    chs::fft_lds_inplace<T>(a, ls, nl, pa.pz, tw, sign, threadIdx.x, blockDim.x);
    for (int t = threadIdx.x; t < nl * N; t += blockDim.x) {
        const int ln = t / N, kz = t - ln * N;
        base[t] = a[ln * ls + kz];
    }
}
template <typename T>
__global__ __launch_bounds__(256) void k_fftz(Geo g, PoisArgs pa, cpx<T> *__restrict__ spec, int NL, int sign)
{
    k_fftz_b<T>(g, pa, spec, NL, sign, own_blk());
}

// Per-mode y-Thomas (TDMAIJJ_nonCYC.f90:84-120 on the FFT99 coefficients,
// possion3d_FFT99.f90:213-231): pp[J][kx][kz] = 1/(b_J - h_{J-1} a_J) precomputed in FP64,
// h_J = c_J pp_J; the (kx,kz) = (0,0) mode has RHS(J=1) = 0 and row 1 decoupled.  P is the
// pivot type of the line (FP64, or FP32 for the well-conditioned modes of the mixed form);
// the arithmetic is FP64 either way.
template <typename T, typename P, int B>
__device__ __forceinline__ void ytdma_line(const int M, const double *__restrict__ amph,
                                           const double *__restrict__ apph, const P *__restrict__ p,
                                           const size_t pst, cpx<T> *__restrict__ s, const size_t st,
                                           const bool zero, const bool lastpin)
{
    // zero: the (0,0) mode with FFT99's gauge (row 1 decoupled, RHS(1) = 0); lastpin: the
    // (0,0) mode with TRID0's (x_M = 0, possion3d_fishpack.f90:322-332), pivots unmodified
    // Rows go in batches of B; the loads of the next batch are issued before the current
    // batch's recurrence (double buffer in registers), and pivots stay in their storage type
    // until used, so no load waits on a conversion.
    cpx<T> rc[B], rn[B];
    P pc[B], pn[B];
    double gr = 0.0, gi = 0.0;
    auto fwd = [&](int J, cpx<T> r, double pj) {
        double rr = (double)r.x, ri = (double)r.y;
        if (zero && J == 1) { rr = 0.0; ri = 0.0; }
        const double a = (J == 1) ? 0.0 : amph[J - 1];
        gr = (rr - a * gr) * pj;
        gi = (ri - a * gi) * pj;
        s[(size_t)(J - 1) * st] = {(T)gr, (T)gi};
    };
    // forward elimination, rows 1..M: full batches, then the remainder row by row
    const int nf = M / B;
    if (nf > 0) {
#pragma unroll
        for (int u = 0; u < B; ++u) {
            rc[u] = s[(size_t)u * st];
            pc[u] = p[(size_t)u * pst];
        }
    }
    for (int b = 0; b < nf; ++b) {
        const int J0 = 1 + b * B;
        const int Jn = (b + 1 < nf) ? J0 + B : J0;   // the last batch re-reads its own rows
#pragma unroll
        for (int u = 0; u < B; ++u) {
            rn[u] = s[(size_t)(Jn + u - 1) * st];
            pn[u] = p[(size_t)(Jn + u - 1) * pst];
        }
#pragma unroll
        for (int u = 0; u < B; ++u) fwd(J0 + u, rc[u], (double)pc[u]);
#pragma unroll
        for (int u = 0; u < B; ++u) {
            rc[u] = rn[u];
            pc[u] = pn[u];
        }
    }
    for (int J = nf * B + 1; J <= M; ++J) fwd(J, s[(size_t)(J - 1) * st], (double)p[(size_t)(J - 1) * pst]);
    // back substitution x_J = g_J - c_J p_J x_{J+1}, rows M-1 down to 1 (x_M = g_M)
    double xr = gr, xi = gi;
    if (lastpin) {
        xr = 0.0;
        xi = 0.0;
        s[(size_t)(M - 1) * st] = {T(0), T(0)};
    }
    auto bwd = [&](int J, cpx<T> gg, double pj) {
        const double c = (zero && J == 1) ? 0.0 : apph[J - 1];
        const double h = c * pj;
        xr = (double)gg.x - h * xr;
        xi = (double)gg.y - h * xi;
        s[(size_t)(J - 1) * st] = {(T)xr, (T)xi};
    };
    const int nb = (M - 1) / B;
    if (nb > 0) {
#pragma unroll
        for (int u = 0; u < B; ++u) {
            rc[u] = s[(size_t)(M - 2 - u) * st];
            pc[u] = p[(size_t)(M - 2 - u) * pst];
        }
    }
    for (int b = 0; b < nb; ++b) {
        const int J0 = M - 1 - b * B;                 // rows J0, J0-1, ..., J0-B+1
        const int Jn = (b + 1 < nb) ? J0 - B : J0;
#pragma unroll
        for (int u = 0; u < B; ++u) {
            rn[u] = s[(size_t)(Jn - u - 1) * st];
            pn[u] = p[(size_t)(Jn - u - 1) * pst];
        }
#pragma unroll
        for (int u = 0; u < B; ++u) bwd(J0 - u, rc[u], (double)pc[u]);
#pragma unroll
        for (int u = 0; u < B; ++u) {
            rc[u] = rn[u];
            pc[u] = pn[u];
        }
    }
    for (int J = M - 1 - nb * B; J >= 1; --J) bwd(J, s[(size_t)(J - 1) * st], (double)p[(size_t)(J - 1) * pst]);
}

// pp00: FISHPACK gauge only, the (0,0) mode's unmodified pivots (M values).
// One thread per (kx, kz) line: (L/2+1) N lines, rows in batches of B.  The lines are few
// (33 K at 256 x 256: half a wavefront per SIMD), so small blocks spread them over every CU
// and longer batches keep more loads in flight per thread.
template <typename T, int B>
__global__ __launch_bounds__(256) void k_ytdma(Geo g, const double *__restrict__ amph, const double *__restrict__ apph,
                                               const double *__restrict__ pp, const float *__restrict__ pp32,
                                               const double *__restrict__ pp00, cpx<T> *__restrict__ spec)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int H = g.H, N = g.N, M = g.M;
    if (t >= H * N) return;
    const int kx = t / N, kz = t - kx * N;
    const bool zero = (kx == 0 && kz == 0);
    const size_t st = (size_t)H * N, off = (size_t)kx * N + kz;
    if (zero && pp00) {
        ytdma_line<T, double, B>(M, amph, apph, pp00, 1, spec, st, false, true);
        return;
    }
    // FP32 pivots for all but the lowest modes (kx <= 1, |kz| <= 1), whose Neumann/near-
    // Neumann y-systems are ill-conditioned on the stretched mesh
    const bool low = !pp32 || (kx <= 1 && (kz <= 1 || kz == N - 1));
    if (low) ytdma_line<T, double, B>(M, amph, apph, pp + off, st, spec + off, st, zero, false);
    else ytdma_line<T, float, B>(M, amph, apph, pp32 + off, st, spec + off, st, zero, false);
}

// The same per-mode y-Thomas with each line split into P chunks of C rows solved side by side,
// so that no intermediate leaves the chip and a line's rows are spread over P wavefronts.
// Block = 64 consecutive lines t = kx N + kz (lanes; a row of the block is 64 consecutive
// spectral values) x P chunks (wavefronts); chunk q holds rows q C + 1 .. min(M, q C + C) of
// its line, and their pivots, in registers.  Both sweeps are linear in what they receive from
// the neighbouring chunk:
//   forward   g_J = gh_J + lam_J g_in,   gh from g_in = 0,  lam_J = (-a_J p_J) lam_{J-1}, lam = 1
//             before the chunk;
//   backward  x_J = xh_J + mu_J x_in,    xh from x_in = 0,  mu_J = (-c_J p_J) mu_{J+1}, mu = 1
//             after the chunk;
// so each chunk publishes its end values (gh, lam at its last row; xh, mu at its first) in LDS,
// composes its neighbours' into g_in / x_in, and corrects its rows.  Chunk 0 receives g_in = 0
// and the last chunk x_in = 0 (x_{M+1} = 0 gives x_M = g_M), so these are the recurrences of
// ytdma_line with the sweeps split by linearity; only rounding differs.  Gauges as ytdma_line.
template <typename T, int C>
__device__ __forceinline__ void k_ytdma_chunk_b(Geo g, const double *__restrict__ amph,
                                                      const double *__restrict__ apph, const double *__restrict__ pp,
                                                      const float *__restrict__ pp32, const double *__restrict__ pp00,
                                                      cpx<T> *__restrict__ spec, int P, const Blk blk)
{
    __shared__ double xr[16][64], xi[16][64], xl[16][64];
    const int HN = g.H * g.N, M = g.M;
    // P: chunks of this launch's lines (blockDim.y on its own launch; an ensemble launch sizes
    // its blocks for its largest M, and the chunks q >= P of a smaller mesh idle)
    const int lane = threadIdx.x, q = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int t0 = blk.x * 64, t = t0 + lane;
    const bool act = t < HN;
    const int tt = act ? t : HN - 1;   // lanes past the last line solve a copy of it, unstored
    const unsigned lo = (unsigned)(tt - t0);   // lane offset from the wavefront's first line
    const int kx = tt / g.N, kz = tt - kx * g.N;
    const bool zero = kx == 0 && kz == 0;
    const bool pin = zero && pp00;      // TRID0: x_M = 0, unmodified pivots pp00[J-1]
    const bool z99 = zero && !pp00;     // FFT99: RHS(1) = 0, row 1 decoupled
    // FP32 pivots (mixed form) for all but the lowest modes, as k_ytdma
    const bool low = !pp32 || (kx <= 1 && (kz <= 1 || kz == g.N - 1));
    const int J0 = q * C + 1, n = min(C, M - q * C);   // rows of this chunk (n >= 1)
    const size_t st = (size_t)HN;
    // row bases are wave-uniform (scalar registers); a lane adds its 32-bit offset lo
    auto srow = [&](int u) { return spec + (size_t)(J0 - 1 + u) * st + t0; };
    double rr[C], ri[C], pv[C];
#pragma unroll
    for (int u = 0; u < C; ++u)
        if (u < n) {
            const cpx<T> v = srow(u)[lo];
            rr[u] = (double)v.x;
            ri[u] = (double)v.y;
            const size_t o = (size_t)(J0 - 1 + u) * st + t0;
            pv[u] = pin ? pp00[J0 - 1 + u] : (low ? (pp + o)[lo] : (double)(pp32 + o)[lo]);
        }
    // forward sweep from g_in = 0
    double gr = 0.0, gi = 0.0, lam = 1.0;
#pragma unroll
    for (int u = 0; u < C; ++u)
        if (u < n) {
            const int J = J0 + u;
            const double a = (J == 1) ? 0.0 : amph[J - 1];
            const bool r0 = z99 && J == 1;
            gr = ((r0 ? 0.0 : rr[u]) - a * gr) * pv[u];
            gi = ((r0 ? 0.0 : ri[u]) - a * gi) * pv[u];
            lam = -(a * pv[u]) * lam;
            rr[u] = gr;
            ri[u] = gi;
        }
    xr[q][lane] = gr;
    xi[q][lane] = gi;
    xl[q][lane] = lam;
    __syncthreads();
    double inr = 0.0, ini = 0.0;
    for (int qq = 0; qq < q; ++qq) {
        inr = xr[qq][lane] + xl[qq][lane] * inr;
        ini = xi[qq][lane] + xl[qq][lane] * ini;
    }
    if (q > 0) {
        lam = 1.0;
#pragma unroll
        for (int u = 0; u < C; ++u)
            if (u < n) {
                lam = -(amph[J0 - 1 + u] * pv[u]) * lam;   // J >= 2 here
                rr[u] += lam * inr;
                ri[u] += lam * ini;
            }
    }
    // backward sweep from x_in = 0
    double xhr = 0.0, xhi = 0.0, mu = 1.0;
#pragma unroll
    for (int u = C - 1; u >= 0; --u)
        if (u < n) {
            const int J = J0 + u;
            const double h = ((z99 && J == 1) ? 0.0 : apph[J - 1]) * pv[u];
            if (pin && J == M) {
                xhr = 0.0;
                xhi = 0.0;
            } else {
                xhr = rr[u] - h * xhr;
                xhi = ri[u] - h * xhi;
            }
            mu = -h * mu;
            rr[u] = xhr;
            ri[u] = xhi;
        }
    __syncthreads();   // every chunk has read the forward end values
    xr[q][lane] = xhr;
    xi[q][lane] = xhi;
    xl[q][lane] = mu;
    __syncthreads();
    double xir = 0.0, xii = 0.0;
    for (int qq = P - 1; qq > q; --qq) {
        xir = xr[qq][lane] + xl[qq][lane] * xir;
        xii = xi[qq][lane] + xl[qq][lane] * xii;
    }
    if (q < P - 1) {
        mu = 1.0;
#pragma unroll
        for (int u = C - 1; u >= 0; --u)
            if (u < n) {
                const int J = J0 + u;
                mu = -(((z99 && J == 1) ? 0.0 : apph[J - 1]) * pv[u]) * mu;
                rr[u] += mu * xir;
                ri[u] += mu * xii;
            }
    }
    if (act) {
#pragma unroll
        for (int u = 0; u < C; ++u)
            if (u < n) srow(u)[lo] = {(T)rr[u], (T)ri[u]};
    }
}
template <typename T, int C>
__global__ __launch_bounds__(1024) void k_ytdma_chunk(Geo g, const double *__restrict__ amph,
                                                      const double *__restrict__ apph, const double *__restrict__ pp,
                                                      const float *__restrict__ pp32, const double *__restrict__ pp00,
                                                      cpx<T> *__restrict__ spec)
{
    k_ytdma_chunk_b<T, C>(g, amph, apph, pp, pp32, pp00, spec, int(blockDim.y), own_blk());
}

// Inverse real x-transform of TK z-planes at row J -> phi (halo field) and the Neumann
// ghost rows of BC_WALL_DPH_tg (BC_WALL.f90:369-386).  Z = A + i B with the Hermitian
// extension; imaginary parts at kx = 0 and L/2 are not used (FFT99.f90:84-86).
template <typename T>
__device__ __forceinline__ void k_fftx_inv_b(Geo g, PoisArgs pa, const cpx<T> *__restrict__ spec,
                                                  double *__restrict__ phi, int TK, const Blk blk)
{
    extern __shared__ unsigned char smem_raw[];
    const int L = g.L, M = g.M, H = g.H, ls = L + 1;
    cpx<T> *tw = reinterpret_cast<cpx<T> *>(smem_raw);
    cpx<T> *a = tw + L;
    const cpx<T> *twg = reinterpret_cast<const cpx<T> *>(pa.twx);
    for (int t = threadIdx.x; t < L; t += blockDim.x) tw[t] = twg[t];
    const int J = blk.x + 1, k0 = blk.y * TK;
    const int np = TK / 2;
    struct C2 { cpx<T> A, B; };
    chs::stage_items<4, C2>(
        H * np,
        [&](int t) {
            const int kx = t / np, p = t - kx * np;
            return C2{spec[six(g, J, kx, k0 + 2 * p)], spec[six(g, J, kx, k0 + 2 * p + 1)]};
        },
        [&](int t, C2 c) {
            const int kx = t / np, p = t - kx * np;
            if (kx == 0 || kx == L / 2) { c.A.y = T(0); c.B.y = T(0); }
            a[p * ls + kx] = {c.A.x - c.B.y, c.A.y + c.B.x};
            if (kx > 0 && kx < L / 2) a[p * ls + L - kx] = {c.A.x + c.B.y, c.B.x - c.A.y};
        });
    __syncthreads();
    chs::fft_lds_inplace<T>(a, ls, np, pa.px, tw, +1, threadIdx.x, blockDim.x);
    for (int t = threadIdx.x; t < np * L; t += blockDim.x) {
        const int p = t / L, i = t - p * L;
        const cpx<T> z = a[p * ls + i];
        for (int h = 0; h < 2; ++h) {
            const int k = k0 + 2 * p + h;
            const double val = h == 0 ? (double)z.x : (double)z.y;
            phi[hix(g, i, J, k)] = val;
            if (J == 1) phi[hix(g, i, 0, k)] = val;
            if (J == M) phi[hix(g, i, M + 1, k)] = val;
        }
    }
}
template <typename T>
__global__ __launch_bounds__(256, 4) void k_fftx_inv(Geo g, PoisArgs pa, const cpx<T> *__restrict__ spec,
                                                  double *__restrict__ phi, int TK)
{
    k_fftx_inv_b<T>(g, pa, spec, phi, TK, own_blk());
}

// ---------------------------------------------------------------------------------------
// Register-FFT forms of the three transform kernels, for line lengths with a plan in
// chs_fftr.h (64 ... 1024).  One wavefront per block owns LPW = 64 / TL whole lines, so the
// passes need only wavefront-level LDS ordering.  Same inputs, outputs, scaling and
// separation as k_divg_fftx / k_fftz / k_fftx_inv.
template <int N> struct RGeom {
    static constexpr int E = chs::RPlan<N>::E, PS = chs::RPlan<N>::PS, TL = N / E, LPW = 64 / TL;
    static constexpr int R0 = chs::rfirst<N>(), LS = chs::rline_stride<N, PS>();
};

// z: lines [J][kx][:] of length N, in place.
template <typename T, int N>
__global__ __launch_bounds__(64) void k_fftz_r(Geo g, const cpx<T> *__restrict__ tw, cpx<T> *__restrict__ spec, int sign)
{
    using G = RGeom<N>;
    __shared__ cpx<T> lds[G::LPW * G::LS];
    const int li = threadIdx.x / G::TL, t = threadIdx.x % G::TL;
    const int nlines = g.M * g.H;
    const int line = blockIdx.x * G::LPW + li;
    const bool act = line < nlines;   // lanes past the last line transform a copy of it, unstored
    cpx<T> *base = spec + (size_t)(act ? line : nlines - 1) * N;
    cpx<T> v[G::E];
#pragma unroll
    for (int b = 0; b < G::E / G::R0; ++b)
#pragma unroll
        for (int r = 0; r < G::R0; ++r) v[b * G::R0 + r] = base[t + b * G::TL + r * (N / G::R0)];
    chs::rfft_line<T, N>(v, lds + li * G::LS, t, tw, sign, [&](int idx, const cpx<T> &x) {
        if (act) base[idx] = x;
    });
}

// x forward: block = row J and TK = 2 LPW consecutive z-planes (plane pair per line).
template <typename T, int L>
__global__ __launch_bounds__(64) void k_divg_fftx_r(Geo g, const cpx<T> *__restrict__ tw, const double *__restrict__ dv,
                                                    double talp,
                                                    const double *__restrict__ q, double *__restrict__ div,
                                                    cpx<T> *__restrict__ spec, const double *__restrict__ src,
                                                    int nxcd)
{
    using G = RGeom<L>;
    constexpr int TK = 2 * G::LPW;
    __shared__ cpx<T> lds[G::LPW * G::LS];
    const int li = threadIdx.x / G::TL, t = threadIdx.x % G::TL;
    // rows J and J+1 (whose v this block also reads) on one XCD, so the re-read hits its L2
    const int tile = __builtin_amdgcn_readfirstlane(xcd_tile(nxcd, own_blk()));
    const int J = tile % gridDim.x + 1, k0 = (tile / gridDim.x) * TK, ka = k0 + 2 * li;
    const size_t hs = (size_t)g.N * (g.M + 2) * L;
    const double *u = q, *v = q + hs, *w = q + 2 * hs;
    const double dyf = g.dyfi[J - 1];
    const double coe0 = dv ? 1.0 / (dv[0] * talp) : 1.0;   // 1 / (alpha DT) (DIVG_tg)
    cpx<T> x[G::E];
#pragma unroll
    for (int b = 0; b < G::E / G::R0; ++b)
#pragma unroll
        for (int r = 0; r < G::R0; ++r) {
            const int i = t + b * G::TL + r * (L / G::R0);
            double d[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int k = ka + h;
                if (src) {
                    d[h] = src[cix(g, i, J, k)];
                } else {
                    const int ip = (i == L - 1) ? 0 : i + 1, kp = (k == g.N - 1) ? 0 : k + 1;
                    const double dd = divg_cell(u[hix(g, ip, J, k)], u[hix(g, i, J, k)], v[hix(g, i, J + 1, k)],
                                                v[hix(g, i, J, k)], w[hix(g, i, J, kp)], w[hix(g, i, J, k)], g.dxi,
                                                dyf, g.dzi);
                    d[h] = dd * coe0;
                    if (div) div[cix(g, i, J, k)] = d[h];
                }
            }
            x[b * G::R0 + r] = {(T)d[0], (T)d[1]};
        }
    cpx<T> *line = lds + li * G::LS;
    chs::rfft_line<T, L>(x, line, t, tw, -1, [&](int idx, const cpx<T> &z) { line[chs::rpad<G::PS>(idx)] = z; });
    chs::wave_lds_sync();
    // separate the two real transforms, scale by 1/(L N), write spec[J][kx][k0 .. k0+TK-1]
    const T scale = T(1.0 / ((double)L * (double)g.N));
    const int ln = threadIdx.x % TK, p = ln >> 1;
    const cpx<T> *zl = lds + p * G::LS;
    for (int kx = threadIdx.x / TK; kx < g.H; kx += 64 / TK) {
        const cpx<T> Z = zl[chs::rpad<G::PS>(kx)];
        const cpx<T> Zc = zl[chs::rpad<G::PS>((L - kx) % L)];
        cpx<T> o;
        if ((ln & 1) == 0) o = {(Z.x + Zc.x) * T(0.5) * scale, (Z.y - Zc.y) * T(0.5) * scale};
        else o = {(Z.y + Zc.y) * T(0.5) * scale, (Zc.x - Z.x) * T(0.5) * scale};
        spec[six(g, J, kx, k0 + ln)] = o;
    }
}

// x inverse: Hermitian extension of the plane pair of each line, complex inverse transform,
// real parts -> phi with the Neumann ghost rows (BC_WALL.f90:369-386).
template <typename T, int L>
__global__ __launch_bounds__(64) void k_fftx_inv_r(Geo g, const cpx<T> *__restrict__ tw,
                                                   const cpx<T> *__restrict__ spec, double *__restrict__ phi)
{
    using G = RGeom<L>;
    constexpr int TK = 2 * G::LPW;
    __shared__ cpx<T> lds[G::LPW * G::LS];
    const int J = blockIdx.x + 1, k0 = blockIdx.y * TK, M = g.M;
    {
        const int p = threadIdx.x % G::LPW;
        cpx<T> *zl = lds + p * G::LS;
        for (int kx = threadIdx.x / G::LPW; kx < g.H; kx += 64 / G::LPW) {
            cpx<T> A = spec[six(g, J, kx, k0 + 2 * p)], B = spec[six(g, J, kx, k0 + 2 * p + 1)];
            if (kx == 0 || kx == L / 2) { A.y = T(0); B.y = T(0); }
            zl[chs::rpad<G::PS>(kx)] = {A.x - B.y, A.y + B.x};
            if (kx > 0 && kx < L / 2) zl[chs::rpad<G::PS>(L - kx)] = {A.x + B.y, B.x - A.y};
        }
    }
    chs::wave_lds_sync();
    const int li = threadIdx.x / G::TL, t = threadIdx.x % G::TL;
    cpx<T> *line = lds + li * G::LS;
    cpx<T> x[G::E];
#pragma unroll
    for (int b = 0; b < G::E / G::R0; ++b)
#pragma unroll
        for (int r = 0; r < G::R0; ++r) x[b * G::R0 + r] = line[chs::rpad<G::PS>(t + b * G::TL + r * (L / G::R0))];
    chs::wave_lds_sync();
    // the block's planes k0 .. k0+TK-1 from one wave-uniform base: a lane's offset is 32-bit
    const auto pb = uniform(phi + (size_t)k0 * (M + 2) * L);
    const unsigned pl = (unsigned)(2 * li) * (unsigned)((M + 2) * L), ps = (unsigned)((M + 2) * L);
    chs::rfft_line<T, L>(x, line, t, tw, +1, [&](int i, const cpx<T> &z) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const unsigned o = pl + h * ps + (unsigned)i;
            const double val = h == 0 ? (double)z.x : (double)z.y;
            pb[o + (unsigned)(J * L)] = val;
            if (J == 1) pb[o] = val;
            if (J == M) pb[o + (unsigned)((M + 1) * L)] = val;
        }
    });
}

// Projection of a substep: VELOUPDT_tg (VELO_CORRECTION.f90:31-71), PRCALC_tg
// (PRESSURE_CORRECTION.f90:26-36) and the ghost rows of BC_WALL_Q_tg / BC_WALL_PR_TG.  The
// predicted velocity qs (written by the y-solve) minus alpha DT grad phi goes to q.  Because the
// prediction has its own buffer, every neighbour is read race-free: the divergence PRCALC_tg
// needs (RHSLLPHI, the Poisson RHS) is recomputed here rather than stored by the transform
// kernel, and with CFL the CFL_tg value of each cell of the corrected field (CFL.f90:27-46) is
// formed from corrected neighbours computed alike, with block maxima into part.  One thread
// per (i, J) marches KZ planes (phi and w* of plane k+1 become the next plane's own).
// Contraction off: a correction is the same product and difference wherever it is formed, so
// a neighbour's corrected value here equals the one its own thread stores.
template <bool CFL>
__device__ __forceinline__ void k_update_b(Geo g, Sub s, const double *__restrict__ phi,
                                                const double *__restrict__ qs, double *__restrict__ q,
                                                double *__restrict__ pr, int KZ, double *__restrict__ part, const Blk blk)
{
#pragma clang fp contract(off)
    const int i = blk.x * 64 + threadIdx.x;
    // a wavefront is one row J: row bases are wave-uniform (scalar), lanes add i, ip or im
    const int J = __builtin_amdgcn_readfirstlane(blk.y * blockDim.y + threadIdx.y + 1);
    const int k0 = blk.z * KZ;
    double cmax = 0.0;
    if (i < g.L && J <= g.M) {
        const int L = g.L, M = g.M, N = g.N;
        const size_t hs = (size_t)N * (M + 2) * L;
        const unsigned ui = i, ip = (i == L - 1) ? 0 : i + 1, im = (i == 0) ? L - 1 : i - 1;
        const int k1 = min(N, k0 + KZ);
        const double sdt = s.dv[0], scv = s.dv[1];
        const double ta = sdt * s.talp;
        const double cx = ta * g.dxi, cz = ta * g.dzi, cy = ta * g.dyci[J];
        const double cyp = J < M ? ta * g.dyci[J + 1] : 0.0;
        const double be = 0.5 * s.talp * sdt * scv;
        const double coe0 = 1.0 / (sdt * s.talp);
        const double dyf = g.dyfi[J - 1];
        auto R = [&](int jj, int kk) { return ((size_t)kk * (M + 2) + jj) * L; };
        auto cr = [&](const double *f, int jj, int kk) { return uniform(f + R(jj, kk)); };
        auto wr = [&](double *f, int jj, int kk) { return uniform(f + R(jj, kk)); };
        double phm = cr(phi, J, k0 == 0 ? N - 1 : k0 - 1)[ui];   // phi at plane k-1
        double phk = cr(phi, J, k0)[ui];
        double wsk = cr(qs + 2 * hs, J, k0)[ui];                 // w* at plane k
        for (int k = k0; k < k1; ++k) {
            const int kp = (k == N - 1) ? 0 : k + 1;
            const auto phr = cr(phi, J, k);
            const double php = cr(phi, J, kp)[ui], wsp = cr(qs + 2 * hs, J, kp)[ui];
            const auto usr = cr(qs, J, k);
            const double us = usr[ui], usp = usr[ip], vs = cr(qs + hs, J, k)[ui], vsp = cr(qs + hs, J + 1, k)[ui];
            const double pim = phr[im], pjm = cr(phi, J - 1, k)[ui];
            const double d = divg_cell(usp, us, vsp, vs, wsp, wsk, g.dxi, dyf, g.dzi) * coe0;
            const double un = us - (phk - pim) * cx;
            const double wn = wsk - (phk - phm) * cz;
            const double vn = J >= 2 ? vs - (phk - pjm) * cy : 0.0;   // row 1 is the bottom wall
            const auto prr = wr(pr, J, k);
            const double pn = prr[ui] + phk - be * d;
            wr(q, J, k)[ui] = un;
            wr(q + hs, J, k)[ui] = vn;
            wr(q + 2 * hs, J, k)[ui] = wn;
            prr[ui] = pn;
            if (J == 1) {
                wr(pr, 0, k)[ui] = pn;
                wr(q, 0, k)[ui] = 0.0;
                wr(q + hs, 0, k)[ui] = 0.0;
                wr(q + 2 * hs, 0, k)[ui] = 0.0;
            }
            if (J == M) {
                wr(pr, M + 1, k)[ui] = pn;
                wr(q, M + 1, k)[ui] = 0.0;
                wr(q + hs, M + 1, k)[ui] = 0.0;
                wr(q + 2 * hs, M + 1, k)[ui] = 0.0;
            }
            if constexpr (CFL) {
                const double unp = usp - (phr[ip] - phk) * cx;
                const double vnp = J < M ? vsp - (cr(phi, J + 1, k)[ui] - phk) * cyp : 0.0;
                const double wnp = wsp - (php - phk) * cz;
                const double a = fabs(un + unp) * g.dxi;
                const double b = fabs(vn + vnp) * dyf;
                const double c = fabs(wn + wnp) * g.dzi;
                cmax = nanmax(cmax, (a + b + c) * 0.5);
            }
            phm = phk;
            phk = php;
            wsk = wsp;
        }
    }
    if constexpr (CFL) {
        __shared__ double red[1024];
        const int t = threadIdx.y * 64 + threadIdx.x;   // blockDim = (64, 2^n)
        red[t] = cmax;
        __syncthreads();
        for (int o = 32 * blockDim.y; o > 0; o >>= 1) {
            if (t < o) red[t] = nanmax(red[t], red[t + o]);
            __syncthreads();
        }
        if (t == 0) part[((size_t)blk.z * blk.gy + blk.y) * blk.gx + blk.x] = red[0];
    }
}
// Tiles dealt to the XCDs in contiguous runs (nxcd > 1), so that the rows a tile reads from
// the tiles above and below it (phi at J-1, v* at J+1) come through the same L2.
template <bool CFL>
__global__ __launch_bounds__(1024) void k_update(Geo g, Sub s, const double *__restrict__ phi,
                                                const double *__restrict__ qs, double *__restrict__ q,
                                                double *__restrict__ pr, int KZ, double *__restrict__ part,
                                                int nxcd)
{
    Blk b = own_blk();
    if (nxcd > 1) {
        const unsigned t = (unsigned)__builtin_amdgcn_readfirstlane(xcd_tile(nxcd, b));
        b.x = t % b.gx;
        b.y = (t / b.gx) % b.gy;
        b.z = t / (b.gx * b.gy);
    }
    k_update_b<CFL>(g, s, phi, qs, q, pr, KZ, part, b);
}

// k_update with two adjacent points per lane (16-byte loads and stores: half the vector-memory
// instructions, which bound k_update): lane l of a wavefront holds x = x0 + 2l and x0 + 2l + 1
// of a 128-point stretch of row J.  The x-neighbours across lanes come by cross-lane moves; the
// stretch's outer ones (phi at x0 - 1, u* and phi at x0 + 128, cyclic) are loaded by the edge
// lanes.  Same arithmetic as k_update_b point by point (contraction off): bit-identical.  L is a
// multiple of 128.
template <bool CFL>
__global__ __launch_bounds__(1024) void k_update2(Geo g, Sub s, const double *__restrict__ phi,
                                                  const double *__restrict__ qs, double *__restrict__ q,
                                                  double *__restrict__ pr, int KZ, double *__restrict__ part)
{
#pragma clang fp contract(off)
    const Blk blk = own_blk();
    const int lane = threadIdx.x;
    const int x0 = blk.x * 128;
    const int J = __builtin_amdgcn_readfirstlane(blk.y * blockDim.y + threadIdx.y + 1);
    const int k0 = blk.z * KZ;
    double cmax = 0.0;
    if (J <= g.M) {
        const int L = g.L, M = g.M, N = g.N;
        const size_t hs = (size_t)N * (M + 2) * L;
        const unsigned xm = x0 == 0 ? L - 1 : x0 - 1, xp = x0 + 128 == L ? 0 : x0 + 128;
        const int k1 = min(N, k0 + KZ);
        const double sdt = s.dv[0], scv = s.dv[1];
        const double ta = sdt * s.talp;
        const double cx = ta * g.dxi, cz = ta * g.dzi, cy = ta * g.dyci[J];
        const double cyp = J < M ? ta * g.dyci[J + 1] : 0.0;
        const double be = 0.5 * s.talp * sdt * scv;
        const double coe0 = 1.0 / (sdt * s.talp);
        const double dyf = g.dyfi[J - 1];
        auto R = [&](int jj, int kk) { return ((size_t)kk * (M + 2) + jj) * L; };
        auto row = [&](const double *f, int jj, int kk) { return uniform(f + R(jj, kk)); };
        typedef double dv2 __attribute__((ext_vector_type(2)));
        auto ld2 = [&](const double *f, int jj, int kk) {
            return reinterpret_cast<const __attribute__((address_space(1))) dv2 *>(row(f, jj, kk) + x0)[lane];
        };
        auto st2 = [&](double *f, int jj, int kk, double a, double b) {
            dv2 v;
            v.x = a;
            v.y = b;
            reinterpret_cast<__attribute__((address_space(1))) dv2 *>(uniform(f + R(jj, kk)) + x0)[lane] = v;
        };
        const auto pm0 = ld2(phi, J, k0 == 0 ? N - 1 : k0 - 1);
        double phm[2] = {pm0.x, pm0.y};
        const auto pk0 = ld2(phi, J, k0), wk0 = ld2(qs + 2 * hs, J, k0);
        double phk[2] = {pk0.x, pk0.y}, wsk[2] = {wk0.x, wk0.y};
        for (int k = k0; k < k1; ++k) {
            const int kp = (k == N - 1) ? 0 : k + 1;
            const auto pp2 = ld2(phi, J, kp), wp2 = ld2(qs + 2 * hs, J, kp);
            const auto u2 = ld2(qs, J, k), v2 = ld2(qs + hs, J, k), vp2 = ld2(qs + hs, J + 1, k);
            const auto pj2 = ld2(phi, J - 1, k), pr2 = ld2(pr, J, k);
            double pjp[2] = {0.0, 0.0};
            if constexpr (CFL) {
                if (J < M) {
                    const auto t = ld2(phi, J + 1, k);
                    pjp[0] = t.x; pjp[1] = t.y;
                }
            }
            // x-neighbours: u*(x+1) and phi(x+1) of the pair's second point from the next lane
            // (the stretch's last lane loads x0 + 128), phi(x-1) of the first from the previous
            double unx = __shfl_down(u2.x, 1), pnx = __shfl_down(phk[0], 1), ppv = __shfl_up(phk[1], 1);
            if (lane == 63) {
                unx = row(qs, J, k)[xp];
                if constexpr (CFL) pnx = row(phi, J, k)[xp];
            }
            if (lane == 0) ppv = row(phi, J, k)[xm];
            const double us[2] = {u2.x, u2.y}, usp[2] = {u2.y, unx};
            const double vs[2] = {v2.x, v2.y}, vsp[2] = {vp2.x, vp2.y};
            const double php[2] = {pp2.x, pp2.y}, wsp[2] = {wp2.x, wp2.y};
            const double pim[2] = {ppv, phk[0]}, pjm[2] = {pj2.x, pj2.y}, pip[2] = {phk[1], pnx};
            const double pold[2] = {pr2.x, pr2.y};
            double un[2], vn[2], wn[2], pn[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const double d = divg_cell(usp[h], us[h], vsp[h], vs[h], wsp[h], wsk[h], g.dxi, dyf, g.dzi) * coe0;
                un[h] = us[h] - (phk[h] - pim[h]) * cx;
                wn[h] = wsk[h] - (phk[h] - phm[h]) * cz;
                vn[h] = J >= 2 ? vs[h] - (phk[h] - pjm[h]) * cy : 0.0;   // row 1 is the bottom wall
                pn[h] = pold[h] + phk[h] - be * d;
                if constexpr (CFL) {
                    const double unp = usp[h] - (pip[h] - phk[h]) * cx;
                    const double vnp = J < M ? vsp[h] - (pjp[h] - phk[h]) * cyp : 0.0;
                    const double wnp = wsp[h] - (php[h] - phk[h]) * cz;
                    const double a = fabs(un[h] + unp) * g.dxi;
                    const double b = fabs(vn[h] + vnp) * dyf;
                    const double c = fabs(wn[h] + wnp) * g.dzi;
                    cmax = nanmax(cmax, (a + b + c) * 0.5);
                }
            }
            st2(q, J, k, un[0], un[1]);
            st2(q + hs, J, k, vn[0], vn[1]);
            st2(q + 2 * hs, J, k, wn[0], wn[1]);
            st2(pr, J, k, pn[0], pn[1]);
            if (J == 1) {
                st2(pr, 0, k, pn[0], pn[1]);
                st2(q, 0, k, 0.0, 0.0);
                st2(q + hs, 0, k, 0.0, 0.0);
                st2(q + 2 * hs, 0, k, 0.0, 0.0);
            }
            if (J == M) {
                st2(pr, M + 1, k, pn[0], pn[1]);
                st2(q, M + 1, k, 0.0, 0.0);
                st2(q + hs, M + 1, k, 0.0, 0.0);
                st2(q + 2 * hs, M + 1, k, 0.0, 0.0);
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                phm[h] = phk[h];
                phk[h] = php[h];
                wsk[h] = wsp[h];
            }
        }
    }
    if constexpr (CFL) {
        __shared__ double red[1024];
        const int t = threadIdx.y * 64 + threadIdx.x;   // blockDim = (64, 2^n)
        red[t] = cmax;
        __syncthreads();
        for (int o = 32 * blockDim.y; o > 0; o >>= 1) {
            if (t < o) red[t] = nanmax(red[t], red[t + o]);
            __syncthreads();
        }
        if (t == 0) part[((size_t)blk.z * blk.gy + blk.y) * blk.gx + blk.x] = red[0];
    }
}

// A refinement sweep's correction, in place: the velocity as k_update, p += phi (PRCALC_tg's
// viscous term used the first solve's RHS only), ghost rows.
__global__ __launch_bounds__(256) void k_update_refine(Geo g, Sub s, const double *__restrict__ phi,
                                                       double *__restrict__ q, double *__restrict__ pr)
{
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int J = blockIdx.y * 4 + threadIdx.y + 1;
    const int k = blockIdx.z;
    if (i >= g.L || J > g.M) return;
    const size_t hs = (size_t)g.N * (g.M + 2) * g.L;
    const int im = (i == 0) ? g.L - 1 : i - 1, km = (k == 0) ? g.N - 1 : k - 1;
    const double p0 = phi[hix(g, i, J, k)];
    const double ta = s.dv[0] * s.talp;
    const double cx = ta * g.dxi, cz = ta * g.dzi;
    const size_t h0 = hix(g, i, J, k);
    q[h0] = q[h0] - (p0 - phi[hix(g, im, J, k)]) * cx;
    q[2 * hs + h0] = q[2 * hs + h0] - (p0 - phi[hix(g, i, J, km)]) * cz;
    if (J >= 2) q[hs + h0] = q[hs + h0] - (p0 - phi[hix(g, i, J - 1, k)]) * (ta * g.dyci[J]);
    const double pn = pr[h0] + p0;
    pr[h0] = pn;
    if (J == 1) {
        pr[hix(g, i, 0, k)] = pn;
        q[hix(g, i, 0, k)] = 0.0;
        q[2 * hs + hix(g, i, 0, k)] = 0.0;
        q[hs + hix(g, i, 0, k)] = 0.0;
        q[hs + h0] = 0.0;
    }
    if (J == g.M) {
        const size_t ht = hix(g, i, g.M + 1, k);
        pr[ht] = pn;
        q[ht] = 0.0;
        q[hs + ht] = 0.0;
        q[2 * hs + ht] = 0.0;
    }
}

// Residual of the discrete Poisson operator the FFT99 solve inverts (possion3d_FFT99.f90:
// 202-231: x and z second differences, y-stencil AMPH/ACPH/APPH with the Neumann folding of
// TDMA_COEF.f90:44-65): r = src - L phi, all FP64, on cell fields; block max |r| partials.
__global__ __launch_bounds__(256) void k_poisson_resid(Geo g, const double *__restrict__ amph,
                                                       const double *__restrict__ acph,
                                                       const double *__restrict__ apph,
                                                       const double *__restrict__ src,
                                                       const double *__restrict__ phi, double *__restrict__ r,
                                                       double *__restrict__ part)
{
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int J = blockIdx.y * 4 + threadIdx.y + 1;
    const int k = blockIdx.z;
    double val = 0.0;
    if (i < g.L && J <= g.M) {
        const int ip = (i == g.L - 1) ? 0 : i + 1, im = (i == 0) ? g.L - 1 : i - 1;
        const int kp = (k == g.N - 1) ? 0 : k + 1, km = (k == 0) ? g.N - 1 : k - 1;
        const double p0 = phi[cix(g, i, J, k)];
        double lap = (phi[cix(g, ip, J, k)] - 2.0 * p0 + phi[cix(g, im, J, k)]) * g.dxqi +
                     (phi[cix(g, i, J, kp)] - 2.0 * p0 + phi[cix(g, i, J, km)]) * g.dzqi + acph[J - 1] * p0;
        if (J > 1) lap += amph[J - 1] * phi[cix(g, i, J - 1, k)];
        if (J < g.M) lap += apph[J - 1] * phi[cix(g, i, J + 1, k)];
        const double rr = src[cix(g, i, J, k)] - lap;
        if (r) r[cix(g, i, J, k)] = rr;
        val = fabs(rr);
    }
    __shared__ double red[256];
    const int t = threadIdx.y * 64 + threadIdx.x;
    red[t] = val;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) red[t] = nanmax(red[t], red[t + o]);
        __syncthreads();
    }
    if (t == 0) part[((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = red[0];
}

// acc (cells) = [acc +] phi interior rows (halo field)
__global__ __launch_bounds__(256) void k_phi_accum(Geo g, const double *__restrict__ phi, double *__restrict__ acc,
                                                   int add, const double *__restrict__ shift)
{
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int J = blockIdx.y * 4 + threadIdx.y + 1;
    const int k = blockIdx.z;
    if (i >= g.L || J > g.M) return;
    double v = phi[hix(g, i, J, k)];
    if (shift) v -= shift[0];
    acc[cix(g, i, J, k)] = add ? acc[cix(g, i, J, k)] + v : v;
}

// Gauge of the pressure correction: the FFT solve fixes phi_hat_00(J=1) = 0, i.e. a zero mean
// over the J = 1 plane (possion3d_FFT99.f90:225-231, :444-446). The residual cannot see a constant, so the
// FP32 rounding of that mean by the first mixed-precision solve survives every refinement sweep;
// out[0] = mean over J = 1 of (acc + correction), which the last accumulate subtracts.
__global__ __launch_bounds__(256) void k_gauge_mean(Geo g, const double *__restrict__ acc,
                                                    const double *__restrict__ phi, double *__restrict__ out)
{
    __shared__ double red[256];
    const int n = g.L * g.N;
    double s = 0.0;
    for (int t = threadIdx.x; t < n; t += 256) {
        const int i = t % g.L, k = t / g.L;
        s += acc[cix(g, i, 1, k)] + phi[hix(g, i, 1, k)];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0] / (double)n;
}

// Diagnostics: CFL_tg (CFL.f90:27-46) max over cells, and DIVGCK_tg max |div| (DIVGCK.f90).
__global__ __launch_bounds__(256) void k_cfl_div(Geo g, const double *__restrict__ q, double *__restrict__ part,
                                                 int which)
{
#pragma clang fp contract(off)   // the reference's rounding: each product and sum rounded
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int J = blockIdx.y * 4 + threadIdx.y + 1;
    const int k = blockIdx.z;
    double val = 0.0;
    if (i < g.L && J <= g.M) {
        const size_t hs = (size_t)g.N * (g.M + 2) * g.L;
        const int ip = (i == g.L - 1) ? 0 : i + 1, kp = (k == g.N - 1) ? 0 : k + 1;
        const double *u = q, *v = q + hs, *w = q + 2 * hs;
        if (which == 0) {
            const double a = fabs(u[hix(g, i, J, k)] + u[hix(g, ip, J, k)]) * g.dxi;
            const double b = fabs(v[hix(g, i, J, k)] + v[hix(g, i, J + 1, k)]) * g.dyfi[J - 1];
            const double c = fabs(w[hix(g, i, J, k)] + w[hix(g, i, J, kp)]) * g.dzi;
            val = (a + b + c) * 0.5;
        } else {
            val = fabs(divg_cell(u[hix(g, ip, J, k)], u[hix(g, i, J, k)], v[hix(g, i, J + 1, k)], v[hix(g, i, J, k)],
                                 w[hix(g, i, J, kp)], w[hix(g, i, J, k)], g.dxi, g.dyfi[J - 1], g.dzi));
        }
    }
    __shared__ double red[256];
    const int t = threadIdx.y * 64 + threadIdx.x;
    red[t] = val;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) red[t] = nanmax(red[t], red[t + o]);
        __syncthreads();
    }
    if (t == 0) part[((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = red[0];
}

// Maximum of n block partials by one block of blockDim.x (a power of two <= 1024) threads.
__device__ __forceinline__ void k_reduce_max_b(const double *__restrict__ part, int n, double *__restrict__ out, const Blk blk)
{
    __shared__ double red[1024];
    double a = 0.0;
    for (int t = threadIdx.x; t < n; t += blockDim.x) a = nanmax(a, part[t]);
    red[threadIdx.x] = a;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] = nanmax(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}
__global__ __launch_bounds__(1024) void k_reduce_max(const double *__restrict__ part, int n, double *__restrict__ out)
{
    k_reduce_max_b(part, n, out, own_blk());
}

// x-z mean statistics, PP_MEAN_ZX_FLOW_TG (processing/PP_MEAN_ZX.f90:14-307), channel.  Per
// row J the 122 statistics are (s): 0..3 U1 (u, v, w, p), 4..6 UP, 7..12 U2 (the reference's
// upper-triangle storage index), 13..22 U3, 23..31 DVDL1 (3 x 3, row major), 32..40 DVDLP,
// 41..121 DVDL2 (9 x 9, row major).  Thread (i, J) sums one group of them over k:
// GRP 0 s = 0..22, GRP 1 s = 23..40, GRP 2..4 the DVDL2 rows l1 = 3 (GRP-2) .. 3 (GRP-2)+2.
template <int GRP>
__global__ __launch_bounds__(64) void k_mean_zx(Geo g, const double *__restrict__ q, const double *__restrict__ pr,
                                                double *__restrict__ part)
{
#pragma clang fp contract(off)
    constexpr int NS = GRP == 0 ? 23 : (GRP == 1 ? 18 : 27);
    constexpr int S0 = GRP == 0 ? 0 : (GRP == 1 ? 23 : 41 + 27 * (GRP - 2));
    const int L = g.L, N = g.N;
    const int i = blockIdx.x * 64 + threadIdx.x, J = blockIdx.y + 1;
    if (i >= L) return;
    const size_t hs = (size_t)N * (g.M + 2) * L;
    const double *U = q, *V = q + hs, *W = q + 2 * hs;
    const int ip = i == L - 1 ? 0 : i + 1, im = i == 0 ? L - 1 : i - 1;
    const double dycp = g.dyci[J + 1], dyc = g.dyci[J], dyf = g.dyfi[J - 1];
    double acc[NS];
#pragma unroll
    for (int t = 0; t < NS; ++t) acc[t] = 0.0;
    for (int k = 0; k < N; ++k) {
        const int kp = k == N - 1 ? 0 : k + 1, km = k == 0 ? N - 1 : k - 1;
        auto X = [&](const double *f, int ii, int jj, int kk) { return f[hix(g, ii, jj, kk)]; };
        double u[3];
        u[0] = X(U, i, J, k) + X(U, ip, J, k);
        u[1] = X(V, i, J, k) + X(V, i, J + 1, k);
        u[2] = X(W, i, J, k) + X(W, i, J, kp);
        const double p = X(pr, i, J, k);
        if constexpr (GRP == 0) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                acc[a] = acc[a] + u[a];
                acc[4 + a] = acc[4 + a] + u[a] * p;
            }
            acc[3] = acc[3] + p;
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = a; b < 3; ++b) {
                    const int l = ((a + 1) * (6 - a)) / 2 + (b + 1) - 3;
                    acc[7 + l - 1] = acc[7 + l - 1] + u[a] * u[b];
                }
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = a; b < 3; ++b)
#pragma unroll
                    for (int c = b; c < 3; ++c) {
                        const int l = (a + 1) * (5 - a) + ((b + 1) * (6 - b)) / 2 + (c + 1) - 8;
                        acc[13 + l - 1] = acc[13 + l - 1] + u[a] * u[b] * u[c];
                    }
        } else {
            const double ujp0 = X(U, i, J + 1, k) + X(U, ip, J + 1, k), ujm0 = X(U, i, J - 1, k) + X(U, ip, J - 1, k);
            const double ukp0 = X(U, i, J, kp) + X(U, ip, J, kp), ukm0 = X(U, i, J, km) + X(U, ip, J, km);
            const double uip2 = X(W, ip, J, k) + X(W, ip, J, kp), uim2 = X(W, im, J, k) + X(W, im, J, kp);
            const double ujp2 = X(W, i, J + 1, k) + X(W, i, J + 1, kp), ujm2 = X(W, i, J - 1, k) + X(W, i, J - 1, kp);
            const double uip1 = X(V, ip, J, k) + X(V, ip, J + 1, k), uim1 = X(V, im, J, k) + X(V, im, J + 1, k);
            const double ukp1 = X(V, i, J, kp) + X(V, i, J + 1, kp), ukm1 = X(V, i, J, km) + X(V, i, J + 1, km);
            double gr[3][3];
            gr[0][0] = X(U, ip, J, k) - X(U, i, J, k);
            gr[0][1] = (ujp0 - u[0]) * dycp + (u[0] - ujm0) * dyc;
            gr[0][2] = ukp0 - ukm0;
            gr[1][0] = uip1 - uim1;
            gr[1][1] = (X(V, i, J + 1, k) - X(V, i, J, k)) * dyf;
            gr[1][2] = ukp1 - ukm1;
            gr[2][0] = uip2 - uim2;
            gr[2][1] = (ujp2 - u[2]) * dycp + (u[2] - ujm2) * dyc;
            gr[2][2] = X(W, i, J, kp) - X(W, i, J, k);
            if constexpr (GRP == 1) {
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int b = 0; b < 3; ++b) {
                        acc[3 * a + b] = acc[3 * a + b] + gr[a][b];
                        acc[9 + 3 * a + b] = acc[9 + 3 * a + b] + gr[a][b] * p;
                    }
            } else {
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    const int l1 = 3 * (GRP - 2) + r;
                    const double gl = gr[l1 / 3][l1 % 3];
#pragma unroll
                    for (int l2 = 0; l2 < 9; ++l2) acc[9 * r + l2] = acc[9 * r + l2] + gl * gr[l2 / 3][l2 % 3];
                }
            }
        }
    }
#pragma unroll
    for (int t = 0; t < NS; ++t) part[((size_t)(J - 1) * 122 + S0 + t) * L + i] = acc[t];
}

// Row sums of the partials in index order, times the statistic's coefficient.
__global__ __launch_bounds__(256) void k_mean_zx_sum(int L, int n, const double *__restrict__ part,
                                                     const double *__restrict__ coef, double *__restrict__ out)
{
#pragma clang fp contract(off)
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    const double *p = part + (size_t)t * L;
    double s = 0.0;
    for (int i = 0; i < L; ++i) s = s + p[i];
    out[t] = s * coef[t % 122];
}

}  // namespace

// =========================================================================================
// Host side
// =========================================================================================
namespace {

thread_local std::string g_err;

bool fft_ok(int n, chs::FftPlan &p)
{
    p.n = n;
    p.npass = 0;
    int m = n;
    while (m % 4 == 0) { p.radix[p.npass++] = 4; m /= 4; }
    while (m % 2 == 0) { p.radix[p.npass++] = 2; m /= 2; }
    while (m % 3 == 0) { p.radix[p.npass++] = 3; m /= 3; }
    while (m % 5 == 0) { p.radix[p.npass++] = 5; m /= 5; }
    return m == 1 && p.npass <= 16;
}

// ---------------------------------------------------------------------------------------
// Launches over many channels (ensembles).  A kind K names a kernel body and its arguments;
// a launch of k_multi<K> runs the body for n items (channels), item i over its own grid
// (gx, gy, gz).  The blocks are numbered over the items in turn (bstart[i] = first block of
// item i); a block finds its item by a binary search (wave-uniform) and runs the body with
// its position in that item's grid.  Every item of a launch has the same block shape and
// dynamic LDS size (the largest any item needs).
template <class A> struct MItem {
    A a;
    unsigned gx, gy, gz;
};

template <class K>
__global__ __launch_bounds__(K::NT, K::MINW) void k_multi(const MItem<typename K::Args> *__restrict__ it,
                                                          const int *__restrict__ bstart, int n)
{
    const int b = blockIdx.x;
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (bstart[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    const MItem<typename K::Args> &m = it[lo];
    const unsigned lb = (unsigned)(b - bstart[lo]), gx = m.gx, gy = m.gy, gz = m.gz;
    K::run(m.a, Blk{lb % gx, (lb / gx) % gy, lb / (gx * gy), gx, gy, gz});
}

template <typename W, bool FENCE> struct KConv {   // k_conv_rhs, 64 x 4 blocks, no x-solve
    struct Args {
        Geo g;
        Sub s;
        const double *q, *pr;
        W *h0, *rhs;
        double *partial;
        int KC;
    };
    static constexpr int NT = 256, MINW = 1;
    static __device__ __forceinline__ void run(const Args &a, Blk blk)
    {
        k_conv_rhs_b<W, 4, 1, FENCE, false>(a.g, a.s, a.q, a.pr, a.h0, a.rhs, a.partial, a.KC, 1, nullptr, 0, blk);
    }
};
struct KReduce {
    struct Args {
        const double *partial;
        int n;
        double scale;
        double *out;
    };
    static constexpr int NT = 256, MINW = 1;
    static __device__ __forceinline__ void run(const Args &a, Blk blk) { k_reduce_b(a.partial, a.n, a.scale, a.out, blk); }
};
template <typename W> struct KCnX {
    struct Args {
        Geo g;
        const W *gx;
        int D, full;
        W *rhs;
        const double *dpg;
        int NLX;
    };
    static constexpr int NT = 256, MINW = 1;
    static __device__ __forceinline__ void run(const Args &a, Blk blk)
    {
        k_cn_x_b<W>(a.g, a.gx, a.D, a.full, a.rhs, a.dpg, a.NLX, blk);
    }
};
template <typename W> struct KCnZ {
    struct Args {
        Geo g;
        const W *gz, *rin;
        W *rout;
        int KZ, D;
        const double *dpg;
    };
    static constexpr int NT = 256, MINW = 1;
    static __device__ __forceinline__ void run(const Args &a, Blk blk)
    {
        switch (a.D) {
#define CHS_MZ(d) case d: k_cn_z_march_b<W, d>(a.g, a.gz, a.rin, a.rout, a.KZ, a.dpg, blk); break;
            CHS_MZ(0) CHS_MZ(1) CHS_MZ(2) CHS_MZ(3) CHS_MZ(4) CHS_MZ(5) CHS_MZ(6) CHS_MZ(7) CHS_MZ(8)
#undef CHS_MZ
        }
    }
};
template <typename W> struct KCnY3 {
    struct Args {
        Geo g;
        const W *by0, *by1, *by2, *rin;
        const double *qin;
        double *qout;
        int KY, D;
    };
    static constexpr int NT = 256, MINW = 1;
    static __device__ __forceinline__ void run(const Args &a, Blk blk)
    {
        switch (a.D) {
#define CHS_MY(d) case d: k_cn_y3_march_b<W, d>(a.g, a.by0, a.by1, a.by2, a.rin, a.qin, a.qout, a.KY, blk); break;
            CHS_MY(0) CHS_MY(1) CHS_MY(2) CHS_MY(3) CHS_MY(4) CHS_MY(5) CHS_MY(6) CHS_MY(7) CHS_MY(8)
#undef CHS_MY
        }
    }
};
template <typename T> struct KDivg {
    struct Args {
        Geo g;
        PoisArgs pa;
        const double *dv;
        double talp;
        const double *q;
        cpx<T> *spec;
        int TK;
    };
    static constexpr int NT = 256, MINW = 4;
    static __device__ __forceinline__ void run(const Args &a, Blk blk)
    {
        k_divg_fftx_b<T>(a.g, a.pa, a.dv, a.talp, a.q, nullptr, a.spec, a.TK, nullptr, blk);
    }
};
template <typename T> struct KFftz {
    struct Args {
        Geo g;
        PoisArgs pa;
        cpx<T> *spec;
        int NL, sign;
    };
    static constexpr int NT = 256, MINW = 1;
    static __device__ __forceinline__ void run(const Args &a, Blk blk) { k_fftz_b<T>(a.g, a.pa, a.spec, a.NL, a.sign, blk); }
};
template <typename T> struct KYtdma {   // C-row chunks; blocks of 64 x (largest chunk count)
    struct Args {
        Geo g;
        const double *amph, *apph, *pp;
        const float *pp32;
        cpx<T> *spec;
        int P, C;
    };
    static constexpr int NT = 1024, MINW = 1;
    static __device__ __forceinline__ void run(const Args &a, Blk blk)
    {
        switch (a.C) {
        case 4: k_ytdma_chunk_b<T, 4>(a.g, a.amph, a.apph, a.pp, a.pp32, nullptr, a.spec, a.P, blk); break;
        case 8: k_ytdma_chunk_b<T, 8>(a.g, a.amph, a.apph, a.pp, a.pp32, nullptr, a.spec, a.P, blk); break;
        default: k_ytdma_chunk_b<T, 16>(a.g, a.amph, a.apph, a.pp, a.pp32, nullptr, a.spec, a.P, blk); break;
        }
    }
};
template <typename T> struct KFftxInv {
    struct Args {
        Geo g;
        PoisArgs pa;
        const cpx<T> *spec;
        double *phi;
        int TK;
    };
    static constexpr int NT = 256, MINW = 4;
    static __device__ __forceinline__ void run(const Args &a, Blk blk) { k_fftx_inv_b<T>(a.g, a.pa, a.spec, a.phi, a.TK, blk); }
};
struct UpdArgs {
    Geo g;
    Sub s;
    const double *phi, *qs;
    double *q, *pr;
    int KZ;
    double *part;
};
template <bool CFL> struct KUpdate {   // 64 x 4 blocks
    using Args = UpdArgs;
    static constexpr int NT = 256, MINW = 1;
    static __device__ __forceinline__ void run(const Args &a, Blk blk)
    {
        k_update_b<CFL>(a.g, a.s, a.phi, a.qs, a.q, a.pr, a.KZ, a.part, blk);
    }
};
struct KReduceMax {
    struct Args {
        const double *part;
        int n;
        double *out;
    };
    static constexpr int NT = 1024, MINW = 1;
    static __device__ __forceinline__ void run(const Args &a, Blk blk) { k_reduce_max_b(a.part, a.n, a.out, blk); }
};
struct KSetStep {
    struct Args {
        double *dv;
        double dt, cvisc;
    };
    static constexpr int NT = 64, MINW = 1;
    static __device__ __forceinline__ void run(const Args &a, Blk blk) { k_set_step_b(a.dv, a.dt, a.cvisc, blk); }
};
template <typename W> struct KCnTables {
    struct Args {
        Geo g;
        CnDev<W> t;
        const double *dv;
        double *stamp;
        int ns0;
    };
    static constexpr int NT = 256, MINW = 1;
    static __device__ __forceinline__ void run(const Args &a, Blk blk) { k_cn_tables_b<W>(a.g, a.t, a.dv, a.stamp, a.ns0, blk); }
};
struct KCnStamp {
    struct Args {
        const double *dv;
        double *stamp;
        int ns0, n;
    };
    static constexpr int NT = 64, MINW = 1;
    static __device__ __forceinline__ void run(const Args &a, Blk blk) { k_cn_stamp_b(a.dv, a.stamp, a.ns0, a.n, blk); }
};

struct CnTab {
    bool valid = false;
    double dt = 0, cvisc = 0;
    int Dx = 0, Dz = 0, Dy[3] = {0, 0, 0};
    int fullx = 0, fullz = 0;
    void *gx = nullptr, *gz = nullptr, *by[3] = {nullptr, nullptr, nullptr};
};

}  // namespace

struct chs_ctx {
    int L = 0, M = 0, N = 0, H = 0, mode = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // host copies (index conventions of Geo)
    std::vector<double> dyfi, dyci, wff, wfb, amvr, acvr, apvr, amph, acph, apph, ak1, ak3;
    double tgam[4], troh[4], talp[4];
    double dxi, dzi, dxqi, dzqi, volm;
    // device
    double *cf = nullptr;   // packed coefficients
    Geo geo;
    const double *d_amph = nullptr, *d_apph = nullptr, *cf_acph = nullptr;
    double *q = nullptr, *pr = nullptr, *phi = nullptr, *div = nullptr;
    double *qs = nullptr;   // predicted velocity of a substep (y-solve output, projection input)
    void *h0 = nullptr, *rhs = nullptr, *rhs2 = nullptr, *spec = nullptr;
    double *pp = nullptr, *partial = nullptr, *scal = nullptr;
    double *dstep = nullptr;   // scal + 8: device step scalars [DT, CVISC, CFLMM, PhyTIME, stop]
    double *cn_stamp = nullptr;   // (DT, CVISC) each stage's Crank-Nicolson tables were built for
    bool dv_host = false;         // dstep holds (dv_dt, dv_cv) set from the host
    bool cfl_valid = false;       // dstep[2] holds CFLMM_tg of the current field
    double dv_dt = 0.0, dv_cv = 0.0;
    float *pp32 = nullptr;   // FP32 copy of the Poisson pivots (mixed form, high modes)
    double *pp00 = nullptr;  // (0,0)-mode pivots without FFT99's decoupling (FISHPACK gauge)
    int npartial = 0;
    void *twx = nullptr, *twz = nullptr;
    void *rtwx = nullptr, *rtwz = nullptr;   // register-FFT pass twiddles (chs_fftr.h), if planned
    chs::FftPlan px, pz;
    CnTab cn[4];
    // launch shapes (chs_set_tuning); 0 = derived from the mesh
    // (defaults: the fastest shapes on MI355X for a 256 x 192 x 256 channel; cnx_nlx 0 means
    // 2048 / L lines)
    struct {
        int conv_by = 4, conv_kc = 16;  // k_conv_rhs rows per block, z-planes per thread
        int conv_bx = 1;                // k_conv_rhs wavefronts along x per block (1 or 4)
        int conv_fence = 0;             // 1: all of a plane's loads stay ahead of its arithmetic
        int cnx_nlx = 0;                // k_cn_x lines per block
        int cnz_kz = 16, cny_ky = 48;   // march lengths of the z and y CN stencils (FP64)
        int fftx_tk = 0, fftz_nl = 0;   // z-planes per x-transform block, z-lines per block
        int nxcd = 1;                   // XCDs the stencil tiles are grouped for (1 = off)
        int fft_reg = 1;                // register FFTs where L / N have a plan (0 = LDS Stockham)
        int cn_xz = 0, cnxz_by = 1;     // fused x+z CN kernel (1; slower on MI355X), rows per block
        int conv_fx = 1;                // x-solve inside k_conv_rhs (rows per block; 0 = off)
        int ytdma_bs = 64, ytdma_rb = 16;   // k_ytdma threads per block, rows per load batch
        int ytdma_c = -1;                   // k_ytdma_chunk rows per chunk (-1 auto = 16, 0 = k_ytdma)
        int conv_lds = 2;                   // k_conv_rhs_lds rows per block (1, 2, 4; 0 = k_conv_rhs)
        int conv_dma = 4;                   // k_conv_rhs_dma rows per block (2, 4; 0 = k_conv_rhs_lds)
        int divg_nxcd = 8;                  // XCDs the k_divg_fftx_r rows are grouped for (1 = off)
        int cny3_ky = 48;                   // k_cn_y3_march rows per chunk (0 = M; -1 = one launch per component)
        int upd_kz = 4;                     // k_update z-planes per thread
        int upd_by = 4;                     // k_update rows per block (1, 2, 4, 8, 16)
        int upd_nxcd = 1;                   // k_update tiles grouped for this many XCDs (1: launch order;
                                            // 8 measured equal at 256x192x256)
        int upd_w2 = 1;                     // k_update2 (two points per lane) where L % 128 == 0 (0 = k_update)
        int cnzy = 1;                       // z and y solves fused (k_cn_zy_march): 1 where L % 64 == 0
                                            // and N % 16 == 0, 2 on every mesh, 0 = two kernels
        int cnzy_ky = 48;                   // k_cn_zy_march rows per chunk (0 = M)
        int cnzy_nxcd = 8;                  // k_cn_zy_march tiles grouped for this many XCDs
    } tune;
    int tune_gen = 0;   // bumped by every chs_set_tuning / chs_set_refinement (graphs are re-made)
    int cn_gen = 0;     // bumped by every rebuild of a Crank-Nicolson table (its memory may move)
    bool have_fields = false;
    // profiling of the dominant kernel
    bool prof = false;
    std::string prof_only;
    // hipGraph of one RK3 step (three substeps), re-captured when dt or cvisc change
    bool use_graph = false;
    int nref_step = 0, nref_projection = 0;   // refinement sweeps (mixed form default 0 / 2)
    bool own_stream = true;   // false for channels of an ensemble (stream owned by chs_ens)
    hipGraphExec_t gexec = nullptr;
    double gdt = 0.0, gcv = 0.0;
    int gtune = -1, gcn = -1;   // tune_gen and cn_gen the step graph was captured with
    int gkind = 0;              // 0: chs_step_tg's graph (host DT), 1: chs_run_tg's (device DT)
    double grule[4] = {0, 0, 0, 0};   // DT0, CFLGV, DTMIN, CVISC of a chs_run_tg graph
    bool dstate = false;        // chs_set_step_state has set DT and PhyTIME
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> ev_used;
};

namespace {

int fail(chs_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    g_err = msg;
    return code;
}

#define HIPCHK(ctx, expr)                                                                  \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(ctx, CHS_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

size_t wsize(const chs_ctx *c) { return c->mode == CHS_MIXED ? sizeof(float) : sizeof(double); }

// Exact inverse of the periodic constant-coefficient tridiagonal (a, b, a), first column.
void circulant_inverse(int n, long double a, long double b, std::vector<long double> &g)
{
    // Solve A g = e0 with the cyclic Thomas algorithm of TDMAIJI_CYC in long double.
    std::vector<long double> H0(n), G1(n), G2(n), r(n, 0.0L);
    r[0] = 1.0L;
    const long double c = a;
    H0[0] = c / b;
    G1[0] = r[0] / b;
    G2[0] = -a / b;
    for (int i = 1; i <= n - 3; ++i) {
        long double pp = 1.0L / (b - H0[i - 1] * a);
        H0[i] = c * pp;
        G1[i] = (r[i] - G1[i - 1] * a) * pp;
        G2[i] = -(G2[i - 1] * a) * pp;
    }
    {
        int i = n - 2;
        long double pp = 1.0L / (b - H0[i - 1] * a);
        G1[i] = (r[i] - G1[i - 1] * a) * pp;
        G2[i] = (-c - G2[i - 1] * a) * pp;
    }
    for (int i = n - 3; i >= 0; --i) {
        G1[i] = -H0[i] * G1[i + 1] + G1[i];
        G2[i] = -H0[i] * G2[i + 1] + G2[i];
    }
    g.assign(n, 0.0L);
    long double xn = (r[n - 1] - c * G1[0] - a * G1[n - 2]) / (b + c * G2[0] + a * G2[n - 2]);
    g[n - 1] = xn;
    for (int i = n - 2; i >= 0; --i) g[i] = G1[i] + G2[i] * xn;
}

// Half-width of the circulant inverse kept so that the dropped entries sum below tol*g0.
int circulant_width(const std::vector<long double> &g, long double tol)
{
    const int n = (int)g.size();
    long double dropped = 0.0L;
    int D = n / 2;
    // entries at cyclic distance d appear at indices d and n-d
    for (int d = n / 2; d >= 1; --d) {
        long double e = fabsl(g[d]) + (d != n - d ? fabsl(g[n - d]) : 0.0L);
        if (dropped + e > tol * fabsl(g[0])) break;
        dropped += e;
        D = d - 1;
    }
    return D;
}

// Exact inverse of the non-periodic tridiagonal (rows j = 0..n-1; a_j sub, b_j diag,
// c_j super), dense, in long double (columns by the Thomas algorithm).
void tridiag_inverse(int n, const std::vector<long double> &a, const std::vector<long double> &b,
                     const std::vector<long double> &c, std::vector<long double> &inv)
{
    inv.assign((size_t)n * n, 0.0L);
    std::vector<long double> H(n), P(n), G(n);
    H[0] = c[0] / b[0];
    P[0] = 1.0L / b[0];
    for (int j = 1; j < n; ++j) {
        P[j] = 1.0L / (b[j] - H[j - 1] * a[j]);
        H[j] = c[j] * P[j];
    }
    for (int col = 0; col < n; ++col) {
        long double gprev = 0.0L;
        for (int j = 0; j < n; ++j) {
            long double r = (j == col) ? 1.0L : 0.0L;
            G[j] = (j == 0) ? r * P[0] : (r - a[j] * gprev) * P[j];
            gprev = G[j];
        }
        long double x = G[n - 1];
        inv[(size_t)(n - 1) * n + col] = x;
        for (int j = n - 2; j >= 0; --j) {
            x = G[j] - H[j] * x;
            inv[(size_t)j * n + col] = x;
        }
    }
}

// Band half-widths of stage ns's Crank-Nicolson inverses at (dt, cvisc), in long double:
// entries whose sum stays below tol of the diagonal (2^-64 FP64, 1e-12 mixed) are dropped.
void cn_widths(const chs_ctx *c, int ns, double dt, double cvisc, CnTab &t)
{
    const bool mixed = c->mode == CHS_MIXED;
    const long double tol = mixed ? 1e-12L : 1e-19L;
    const long double beta = 0.5L * (long double)c->talp[ns] * (long double)dt * (long double)cvisc;
    {   // x (MOMFA1_X_tg, MOMFA_tg.f90:72-82): a = c = -FACOE*DXQI, b = 1 - a - c
        const long double co = -beta * (long double)c->dxqi;
        std::vector<long double> g;
        circulant_inverse(c->L, co, 1.0L - co - co, g);
        t.Dx = circulant_width(g, tol);
        t.fullx = (2 * t.Dx + 1 >= c->L) ? 1 : 0;
    }
    {   // z (MOMFA2_Z_tg, MOMFA_tg.f90:129-146)
        const long double co = -beta * (long double)c->dzqi;
        std::vector<long double> g;
        circulant_inverse(c->N, co, 1.0L - co - co, g);
        t.Dz = circulant_width(g, tol);
        t.fullz = (2 * t.Dz + 1 >= c->N) ? 1 : 0;
    }
    const int M = c->M;   // y (MOMFA3_Y_tg, MOMFA_tg.f90:206-216): rows J = 1..NCL2
    for (int comp = 0; comp < 3; ++comp) {
        std::vector<long double> a(M), b(M), cc(M), inv;
        for (int J = 1; J <= M; ++J) {
            const size_t o = (size_t)comp * (M + 2) + J;
            b[J - 1] = 1.0L - beta * (long double)c->acvr[o];
            cc[J - 1] = -beta * (long double)c->apvr[o];
            a[J - 1] = -beta * (long double)c->amvr[o];
        }
        a[0] = 0.0L;       // multiplies BCJ(:,1) = 0 (TDMAIJJ_nonCYC.f90:87)
        cc[M - 1] = 0.0L;  // multiplies BCJ(:,2) = 0 (TDMAIJJ_nonCYC.f90:102)
        tridiag_inverse(M, a, b, cc, inv);
        int D = 0;
        for (int J = 0; J < M; ++J) {
            const long double diag = fabsl(inv[(size_t)J * M + J]);
            long double dropped = 0.0L;
            int dJ = M - 1;
            for (int d = M - 1; d >= 1; --d) {
                long double e = 0.0L;
                if (J + d < M) e += fabsl(inv[(size_t)J * M + J + d]);
                if (J - d >= 0) e += fabsl(inv[(size_t)J * M + J - d]);
                if (dropped + e > tol * diag) break;
                dropped += e;
                dJ = d - 1;
            }
            if (dJ > D) D = dJ;
        }
        t.Dy[comp] = D;
    }
    // the three components share the widest half-width (entries past a component's own width
    // are below the tolerance), so that one launch solves all three with one window
    const int Dm = std::max(t.Dy[0], std::max(t.Dy[1], t.Dy[2]));
    for (int comp = 0; comp < 3; ++comp) t.Dy[comp] = Dm;
}

// Device tables of every stage, at full size once (x: L entries, z: N, y: M rows of up to
// 2M - 1), so that no rebuild moves them.
int cn_alloc(chs_ctx *c)
{
    if (c->cn[0].gx) return 0;
    const size_t ws = wsize(c), M = c->M;
    for (auto &t : c->cn) {
        HIPCHK(c, hipMalloc(&t.gx, ws * c->L));
        HIPCHK(c, hipMalloc(&t.gz, ws * c->N));
        for (auto &b : t.by) HIPCHK(c, hipMalloc(&b, ws * M * (2 * M - 1)));
    }
    HIPCHK(c, hipMalloc(&c->cn_stamp, sizeof(double) * 8));
    // stamps no DT can equal: the first launch of k_cn_tables builds every stage
    const double nan = std::nan("");
    const double st[8] = {nan, nan, nan, nan, nan, nan, nan, nan};
    HIPCHK(c, hipMemcpy(c->cn_stamp, st, sizeof st, hipMemcpyHostToDevice));
    return 0;
}

template <typename W>
CnDev<W> cn_dev(const chs_ctx *c)
{
    CnDev<W> d{};
    for (int ns = 0; ns < 4; ++ns) {
        d.gx[ns] = (W *)c->cn[ns].gx;
        d.gz[ns] = (W *)c->cn[ns].gz;
        for (int k = 0; k < 3; ++k) {
            d.by[ns][k] = (W *)c->cn[ns].by[k];
            d.Dy[ns][k] = c->cn[ns].Dy[k];
        }
        d.talp[ns] = c->talp[ns];
    }
    return d;
}

// Enqueues the rebuild of stages ns0 .. ns0 + n - 1 for the device's current DT and CVISC.
void cn_launch(chs_ctx *c, int ns0, int n)
{
    const dim3 grid(5, n);
    if (c->mode == CHS_MIXED)
        hipLaunchKernelGGL(k_cn_tables<float>, grid, dim3(256), 0, c->stream, c->geo, cn_dev<float>(c), c->dstep,
                           c->cn_stamp, ns0);
    else
        hipLaunchKernelGGL(k_cn_tables<double>, grid, dim3(256), 0, c->stream, c->geo, cn_dev<double>(c), c->dstep,
                           c->cn_stamp, ns0);
    hipLaunchKernelGGL(k_cn_stamp, dim3(1), dim3(64), 0, c->stream, c->dstep, c->cn_stamp, ns0, n);
}

// Host-given DT and CVISC for the next launches (a no-op when the device already holds them).
void set_step(chs_ctx *c, double dt, double cvisc)
{
    if (c->dv_host && c->dv_dt == dt && c->dv_cv == cvisc) return;
    hipLaunchKernelGGL(k_set_step, dim3(1), dim3(64), 0, c->stream, c->dstep, dt, cvisc);
    c->dv_host = true;
    c->dv_dt = dt;
    c->dv_cv = cvisc;
}

// Stage ns's tables for a host-given (dt, cvisc): widths on the host, entries on the device
// (no synchronisation).  A change of widths changes the kernels a step launches, so captured
// graphs are re-made (cn_gen).
int build_cn(chs_ctx *c, int ns, double dt, double cvisc)
{
    CnTab &t = c->cn[ns];
    if (t.valid && t.dt == dt && t.cvisc == cvisc) return 0;
    if (int e = cn_alloc(c)) return e;
    CnTab w = t;
    cn_widths(c, ns, dt, cvisc, w);
    const bool changed = w.Dx != t.Dx || w.Dz != t.Dz || w.fullx != t.fullx || w.fullz != t.fullz ||
                         w.Dy[0] != t.Dy[0] || w.Dy[1] != t.Dy[1] || w.Dy[2] != t.Dy[2];
    if (!t.valid || changed) ++c->cn_gen;
    t.Dx = w.Dx; t.Dz = w.Dz; t.fullx = w.fullx; t.fullz = w.fullz;
    for (int k = 0; k < 3; ++k) t.Dy[k] = w.Dy[k];
    if (changed) hipLaunchKernelGGL(k_cn_stamp, dim3(1), dim3(64), 0, c->stream, nullptr, c->cn_stamp, ns, 1);
    set_step(c, dt, cvisc);
    cn_launch(c, ns, 1);
    HIPCHK(c, hipGetLastError());
    t.dt = dt;
    t.cvisc = cvisc;
    t.valid = true;
    return 0;
}

template <typename T>
int upload_twiddles(chs_ctx *c, int n, void **dst)
{
    std::vector<cpx<T>> h(n);
    for (int m = 0; m < n; ++m) {
        long double ang = -2.0L * 3.141592653589793238462643383279502884L * (long double)m / (long double)n;
        h[m] = {(T)cosl(ang), (T)sinl(ang)};
    }
    HIPCHK(c, hipMalloc(dst, sizeof(cpx<T>) * n));
    HIPCHK(c, hipMemcpyAsync(*dst, h.data(), sizeof(cpx<T>) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

template <typename T>
int upload_rtw(chs_ctx *c, int n, void **dst)
{
    const std::vector<cpx<T>> h = chs::rplan_twiddles<T>(n);
    HIPCHK(c, hipMalloc(dst, sizeof(cpx<T>) * h.size()));
    HIPCHK(c, hipMemcpyAsync(*dst, h.data(), sizeof(cpx<T>) * h.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

// z-planes per x-transform block: TK x-lines = TK/2 complex lines, (TK/2)*L <= 2048 complex
// values (the in-place FFT's register budget), TK a power of two dividing N, at most 16
// (16 complex values = 256 contiguous bytes per kx on the spectral write).
int tk_for(int L, int N)
{
    int tk = 16;
    while (tk > 2 && (N % tk != 0 || (tk / 2) * L > 2048)) tk >>= 1;
    return tk;
}
// contiguous z-lines per z-transform block
int nlz_for(int N) { int n = 2048 / N; return n < 1 ? 1 : n; }


struct EvScope {
    chs_ctx *c;
    const char *name;
    hipEvent_t e1 = nullptr;
    EvScope(chs_ctx *cc, const char *nm) : c(cc), name(nm)
    {
        if (!c->prof || (!c->prof_only.empty() && c->prof_only != nm)) return;
        (void)hipEventCreate(&e1);
        (void)hipEventRecord(e1, c->stream);
    }
    ~EvScope()
    {
        if (!c->prof || !e1) return;
        hipEvent_t e2;
        (void)hipEventCreate(&e2);
        (void)hipEventRecord(e2, c->stream);
        c->ev_used.push_back({name, {e1, e2}});
    }
};

#define CHS_RCASES(X) X(64) X(128) X(256) X(512) X(1024)
template <typename T>
void fftz_r(int N, dim3 grid, hipStream_t st, const Geo &g, const cpx<T> *tw, cpx<T> *spec, int sign)
{
    switch (N) {
#define X(n) case n: hipLaunchKernelGGL((k_fftz_r<T, n>), grid, dim3(64), 0, st, g, tw, spec, sign); break;
        CHS_RCASES(X)
#undef X
    }
}
template <typename T>
void divg_fftx_r(int L, dim3 grid, hipStream_t st, const Geo &g, const cpx<T> *tw, const double *dv, double talp,
                 const double *q, double *div, cpx<T> *spec, const double *src, int nxcd)
{
    switch (L) {
#define X(n) case n: hipLaunchKernelGGL((k_divg_fftx_r<T, n>), grid, dim3(64), 0, st, g, tw, dv, talp, q, div, spec, src, nxcd); break;
        CHS_RCASES(X)
#undef X
    }
}
template <typename T>
void fftx_inv_r(int L, dim3 grid, hipStream_t st, const Geo &g, const cpx<T> *tw, const cpx<T> *spec, double *phi)
{
    switch (L) {
#define X(n) case n: hipLaunchKernelGGL((k_fftx_inv_r<T, n>), grid, dim3(64), 0, st, g, tw, spec, phi); break;
        CHS_RCASES(X)
#undef X
    }
}
#undef CHS_RCASES

// Rows per chunk of k_ytdma_chunk (4, 8 or 16; at most 16 chunks, so M <= 256 — longer
// chunks exceed the 128 registers a 1024-thread block allows): the tuning value if it fits;
// automatic (-1): 16, or on meshes with fewer than 256 blocks of 64 lines the shortest
// chunks that fit (more wavefronts per block where blocks are too few to fill the GPU);
// 0, or a mesh with M > 256, selects the one-thread-per-line k_ytdma.
int ytdma_chunk_rows(int want, int M, int lines)
{
    if (want == 0 || M > 16 * 16) return 0;
    if ((want == 4 || want == 8) && (M + want - 1) / want <= 16) return want;
    if (want < 0 && (lines + 63) / 64 < 256)
        for (int c : {4, 8})
            if ((M + c - 1) / c <= 16) return c;
    return 16;
}

// gauge 0: FFT99's (0,0) rule (row 1 decoupled), 1: TRID0's (last row pinned, FISHPACK)
template <typename T>
// dv: device DT for the divergence scaling 1 / (talp DT) of a substep (null: a given RHS, src).
// The RHS is the divergence of the velocity at vel (c->q when null), stored into div (c->div
// when null; not stored when div_out is false), or the cell field src.
void poisson_launch(chs_ctx *c, const double *dv, double talp, const double *src, int gauge = 0,
                    const double *vel = nullptr, bool div_out = true)
{
    const double *qv = vel ? vel : c->q;
    double *dvo = div_out ? c->div : nullptr;
    const Geo &g = c->geo;
    const int L = c->L, M = c->M, N = c->N, H = c->H;
    const bool rx = c->tune.fft_reg && c->rtwx, rz = c->tune.fft_reg && c->rtwz;
    const int TKR = 2 * (rx ? chs::rplan_lpw(L) : 1), LPZ = rz ? chs::rplan_lpw(N) : 1;
    PoisArgs pa{c->px, c->pz, c->twx, c->twz};
    const int TK = c->tune.fftx_tk > 0 ? c->tune.fftx_tk : tk_for(L, N);
    const int NL = c->tune.fftz_nl > 0 ? c->tune.fftz_nl : nlz_for(N);
    const size_t shx = sizeof(cpx<T>) * (L + (TK / 2) * (L + 1));
    const size_t shz = sizeof(cpx<T>) * (N + NL * (N + 1));
    cpx<T> *spec = (cpx<T> *)c->spec;
    {
        EvScope ev(c, "divg_fftx");
        if (rx)
            divg_fftx_r<T>(L, dim3(M, N / TKR), c->stream, g, (const cpx<T> *)c->rtwx, dv, talp, qv, dvo, spec, src,
                           c->tune.divg_nxcd);
        else
            hipLaunchKernelGGL(k_divg_fftx<T>, dim3(M, N / TK), dim3(256), shx, c->stream, g, pa, dv, talp, qv, dvo,
                               spec, TK, src);
    }
    {
        EvScope ev(c, "fftz_fwd");
        if (rz)
            fftz_r<T>(N, dim3((M * H + LPZ - 1) / LPZ), c->stream, g, (const cpx<T> *)c->rtwz, spec, -1);
        else
            hipLaunchKernelGGL(k_fftz<T>, dim3((M * H + NL - 1) / NL), dim3(256), shz, c->stream, g, pa, spec, NL, -1);
    }
    {
        EvScope ev(c, "ytdma");
        const float *p32 = c->mode == CHS_MIXED ? c->pp32 : (const float *)nullptr;
        const double *p00 = gauge ? (const double *)c->pp00 : (const double *)nullptr;
        const int C = ytdma_chunk_rows(c->tune.ytdma_c, M, H * N);
        if (C > 0) {
            const dim3 grid((H * N + 63) / 64), blk(64, (M + C - 1) / C);
            switch (C) {
#define CHS_YTC(cc) case cc: hipLaunchKernelGGL((k_ytdma_chunk<T, cc>), grid, blk, 0, c->stream, g, c->d_amph, c->d_apph, c->pp, p32, p00, spec); break;
                CHS_YTC(4) CHS_YTC(8) CHS_YTC(16)
#undef CHS_YTC
            }
        } else {
            const int bs = c->tune.ytdma_bs;
            auto kern = c->tune.ytdma_rb == 8 ? k_ytdma<T, 8> : k_ytdma<T, 16>;
            hipLaunchKernelGGL(kern, dim3((H * N + bs - 1) / bs), dim3(bs), 0, c->stream, g, c->d_amph,
                               c->d_apph, c->pp, p32, p00, spec);
        }
    }
    {
        EvScope ev(c, "fftz_inv");
        if (rz)
            fftz_r<T>(N, dim3((M * H + LPZ - 1) / LPZ), c->stream, g, (const cpx<T> *)c->rtwz, spec, +1);
        else
            hipLaunchKernelGGL(k_fftz<T>, dim3((M * H + NL - 1) / NL), dim3(256), shz, c->stream, g, pa, spec, NL, +1);
    }
    {
        EvScope ev(c, "fftx_inv");
        if (rx)
            fftx_inv_r<T>(L, dim3(M, N / TKR), c->stream, g, (const cpx<T> *)c->rtwx, (const cpx<T> *)spec, c->phi);
        else
            hipLaunchKernelGGL(k_fftx_inv<T>, dim3(M, N / TK), dim3(256), shx, c->stream, g, pa,
                               (const cpx<T> *)spec, c->phi, TK);
    }
}

template <typename W>
void cn_z_dispatch(int D, dim3 grid, dim3 blk, hipStream_t st, const Geo &g, const W *gz, const W *rin, W *rout, int KZ,
                   const double *dpg)
{
    switch (D) {
#define CHS_ZCASE(d) case d: hipLaunchKernelGGL((k_cn_z_march<W, d>), grid, blk, 0, st, g, gz, rin, rout, KZ, dpg); break;
        CHS_ZCASE(0) CHS_ZCASE(1) CHS_ZCASE(2) CHS_ZCASE(3) CHS_ZCASE(4) CHS_ZCASE(5) CHS_ZCASE(6)
        CHS_ZCASE(7) CHS_ZCASE(8)
#undef CHS_ZCASE
    }
}

template <typename W>
void cn_xz_dispatch(int D, dim3 grid, dim3 blk, size_t sh, hipStream_t st, const Geo &g, const W *gx, int Dx,
                    int fullx, const W *gz, const W *rin, W *rout, const double *dpg, int KZ)
{
    switch (D) {
#define CHS_XZCASE(d) case d: hipLaunchKernelGGL((k_cn_xz<W, d>), grid, blk, sh, st, g, gx, Dx, fullx, gz, rin, rout, dpg, KZ); break;
        CHS_XZCASE(0) CHS_XZCASE(1) CHS_XZCASE(2) CHS_XZCASE(3) CHS_XZCASE(4) CHS_XZCASE(5) CHS_XZCASE(6)
        CHS_XZCASE(7) CHS_XZCASE(8)
#undef CHS_XZCASE
    }
}

template <typename W>
void cn_y_dispatch(int D, dim3 grid, dim3 blk, hipStream_t st, const Geo &g, const W *by, W *rin, const double *qin,
                   double *qout, int comp, int KY)
{
    switch (D) {
#define CHS_YCASE(d) case d: hipLaunchKernelGGL((k_cn_y_march<W, d>), grid, blk, 0, st, g, by, rin, qin, qout, comp, KY); break;
        CHS_YCASE(0) CHS_YCASE(1) CHS_YCASE(2) CHS_YCASE(3) CHS_YCASE(4) CHS_YCASE(5) CHS_YCASE(6)
        CHS_YCASE(7) CHS_YCASE(8)
#undef CHS_YCASE
    }
}

// One substep SOLVERRK3_MOM_tg(ns).  host_dt: DT and CVISC given by the host (tables and the
// device step scalars are brought to them first); otherwise the device already holds them
// (a step whose DT the device chose).  cfl: the correction also forms CFLMM_tg of the new
// field into dstep[2] (CFL_tg, CFL.f90:13-57).
template <typename W>
int substep_impl(chs_ctx *c, int ns, double dt, double cvisc, bool host_dt = true, bool cfl = false)
{
    if (host_dt) {
        if (int e = build_cn(c, ns, dt, cvisc)) return e;
        set_step(c, dt, cvisc);
    }
    const Geo &g = c->geo;
    const Sub s = {c->tgam[ns], c->troh[ns], c->talp[ns], c->dstep};
    const int L = c->L, M = c->M, N = c->N;
    const size_t cs = (size_t)N * M * L;
    W *h0 = (W *)c->h0, *rhs = (W *)c->rhs, *rhs2 = (W *)c->rhs2;
    const dim3 blk(64, 4), grd((L + 63) / 64, (M + 3) / 4, N);
    const int BY = c->tune.conv_by, KC = c->tune.conv_kc, BX = c->tune.conv_bx;
    CnTab &t = c->cn[ns];
    const int KZ = c->tune.cnz_kz, KY = c->tune.cny_ky;
    const bool zmarch = !t.fullz && t.Dz <= 8;
    const int LX = 64 * ((L + 63) / 64), BYX = c->tune.cnxz_by;
    const size_t shxz = 2 * 8 * (size_t)BYX * LX * sizeof(W);
    const bool xz = c->tune.cn_xz && zmarch && LX * BYX <= 1024 && shxz <= 65536;
    // x-solve inside k_conv_rhs: whole x-lines of 1..4 wavefronts, a short x-inverse band, and
    // the z-marching kernel to take the deferred mass-flux correction
    const int FY = c->tune.conv_fx;   // rows per block of the fused kernel (0: not fused)
    const bool fx = FY > 0 && M % FY == 0 && !xz && zmarch && L % 64 == 0 && L * FY <= 512 && !t.fullx &&
                    t.Dx <= 16;
    // state staged through LDS (k_conv_rhs_lds): LY rows of whole x-lines per block
    const int LY = c->tune.conv_lds;
    // state planes moved into LDS by direct loads (k_conv_rhs_dma): DY rows per block
    const int DY = c->tune.conv_dma;
    const bool dma = fx && DY > 0 && M % DY == 0 && L <= 256 && L * DY <= 1024;
    const bool lds = !dma && fx && LY > 0 && M % LY == 0 && L <= 256 && L * LY <= 1024;
    const int RY = dma ? DY : LY;
    const dim3 blkc = (dma || lds) ? dim3(L, RY) : fx ? dim3(L, FY) : dim3(64 * BX, BY);
    const dim3 grdc = (dma || lds) ? dim3(1, M / RY, (N + KC - 1) / KC)
                      : fx         ? dim3(1, M / FY, (N + KC - 1) / KC)
                                   : dim3((L + 64 * BX - 1) / (64 * BX), (M + BY - 1) / BY, (N + KC - 1) / KC);
    const int npart = grdc.x * grdc.y * grdc.z;
    if (dma) {
        EvScope ev(c, "conv_rhs");
        switch (DY * 10 + L / 64) {
#define CHS_DCASE(by, bx) case by * 10 + bx: hipLaunchKernelGGL((k_conv_rhs_dma<W, by, bx>), grdc, blkc, 0, c->stream, g, s, c->q, c->pr, h0, rhs, c->partial, KC, c->tune.nxcd, (const W *)t.gx, t.Dx); break;
            CHS_DCASE(2, 1) CHS_DCASE(2, 2) CHS_DCASE(2, 3) CHS_DCASE(2, 4)
            CHS_DCASE(4, 1) CHS_DCASE(4, 2) CHS_DCASE(4, 3) CHS_DCASE(4, 4)
#undef CHS_DCASE
        }
    } else if (lds) {
        EvScope ev(c, "conv_rhs");
        switch (LY * 10 + L / 64) {
#define CHS_LCASE(by, bx) case by * 10 + bx: hipLaunchKernelGGL((k_conv_rhs_lds<W, by, bx>), grdc, blkc, 0, c->stream, g, s, c->q, c->pr, h0, rhs, c->partial, KC, c->tune.nxcd, (const W *)t.gx, t.Dx); break;
            CHS_LCASE(1, 1) CHS_LCASE(1, 2) CHS_LCASE(1, 3) CHS_LCASE(1, 4)
            CHS_LCASE(2, 1) CHS_LCASE(2, 2) CHS_LCASE(2, 3) CHS_LCASE(2, 4)
            CHS_LCASE(4, 1) CHS_LCASE(4, 2) CHS_LCASE(4, 3) CHS_LCASE(4, 4)
#undef CHS_LCASE
        }
    } else if (fx) {
        EvScope ev(c, "conv_rhs");
        const int fence = c->tune.conv_fence;
        switch (fence * 100 + FY * 10 + L / 64) {
#define CHS_FCASE(f, by, bx) case f * 100 + by * 10 + bx: hipLaunchKernelGGL((k_conv_rhs<W, by, bx, f, true>), grdc, blkc, 0, c->stream, g, s, c->q, c->pr, h0, rhs, c->partial, KC, c->tune.nxcd, (const W *)t.gx, t.Dx); break;
            CHS_FCASE(0, 1, 1) CHS_FCASE(0, 1, 2) CHS_FCASE(0, 1, 3) CHS_FCASE(0, 1, 4)
            CHS_FCASE(0, 2, 1) CHS_FCASE(0, 2, 2) CHS_FCASE(0, 2, 3) CHS_FCASE(0, 2, 4)
            CHS_FCASE(1, 1, 1) CHS_FCASE(1, 1, 2) CHS_FCASE(1, 1, 3) CHS_FCASE(1, 1, 4)
            CHS_FCASE(1, 2, 1) CHS_FCASE(1, 2, 2) CHS_FCASE(1, 2, 3) CHS_FCASE(1, 2, 4)
#undef CHS_FCASE
        }
    } else {
        EvScope ev(c, "conv_rhs");
        switch (c->tune.conv_fence * 1000 + BX * 100 + BY) {
#define CHS_CCASE(f, bx, by) case f * 1000 + bx * 100 + by: hipLaunchKernelGGL((k_conv_rhs<W, by, bx, f>), grdc, blkc, 0, c->stream, g, s, c->q, c->pr, h0, rhs, c->partial, KC, c->tune.nxcd); break;
            CHS_CCASE(0, 1, 1) CHS_CCASE(0, 1, 2) CHS_CCASE(0, 1, 4) CHS_CCASE(0, 1, 8) CHS_CCASE(0, 1, 16)
            CHS_CCASE(0, 4, 1) CHS_CCASE(0, 4, 2) CHS_CCASE(0, 4, 4)
            CHS_CCASE(1, 1, 1) CHS_CCASE(1, 1, 2) CHS_CCASE(1, 1, 4) CHS_CCASE(1, 1, 8) CHS_CCASE(1, 1, 16)
            CHS_CCASE(1, 4, 1) CHS_CCASE(1, 4, 2) CHS_CCASE(1, 4, 4)
#undef CHS_CCASE
        }
    }
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(256), 0, c->stream, c->partial, npart, 1.0 / c->volm, c->scal);
    // z then y CN stencils: marching kernels for band half-widths up to 8 (the configs need
    // 3..7), the direct-load kernels otherwise; x and z fused where a block holds a whole
    // x-line (L <= 1024) and its two batches of planes fit 64 KiB of LDS
    if (fx) {
        // the x-solve ran inside k_conv_rhs
    } else if (xz) {
        EvScope ev(c, "cn_xz");
        const dim3 gxz((M + BYX - 1) / BYX, 3 * ((N + KZ - 1) / KZ));
        cn_xz_dispatch<W>(t.Dz, gxz, dim3(LX, BYX), shxz, c->stream, g, (const W *)t.gx, t.Dx, t.fullx,
                          (const W *)t.gz, rhs, rhs2, c->scal, KZ);
    } else {
        const int NLX = c->tune.cnx_nlx > 0 ? c->tune.cnx_nlx : ((2048 / L) < 1 ? 1 : 2048 / L);
        EvScope ev(c, "cn_x");
        hipLaunchKernelGGL(k_cn_x<W>, dim3((M * N + NLX - 1) / NLX, 3), dim3(256), sizeof(W) * NLX * L, c->stream,
                           g, (const W *)t.gx, t.Dx, t.fullx, rhs, c->scal, NLX);
    }
    const int Dym = t.Dy[0];
    const bool ysame = t.Dy[1] == Dym && t.Dy[2] == Dym;
    // z and y fused (k_cn_zy_march): wavefronts of 64 x, blocks of 16 planes, z-halo within
    // the block's side planes; batches of 2 FP64 / 4 FP32 rows, so that two blocks fit a CU
    // (48 KB of LDS each at Dz <= 4, and at most 64 VGPRs)
    constexpr int ZYKB = 16, ZYBT = sizeof(W) == 8 ? 2 : 4;
    const int KZY = c->tune.cnzy_ky > 0 ? c->tune.cnzy_ky : M;
    // (by default only on whole tiles: on the ensemble's uneven meshes the ragged tiles made the
    // pass slower than the two kernels)
    const bool full = L % 64 == 0 && N % ZYKB == 0;
    const bool zy = (c->tune.cnzy == 2 || (c->tune.cnzy == 1 && full)) && !xz && zmarch && 2 * t.Dz <= ZYKB &&
                    t.Dz <= 8 && ysame && Dym <= 8 && KZY > Dym;
    if (zy) {
        EvScope ev(c, "cn_zy");
        const int nb = ((L + 63) / 64) * ((N + ZYKB - 1) / ZYKB) * 3 * ((M + KZY - 1) / KZY);
        const W *b0 = (const W *)t.by[0], *b1 = (const W *)t.by[1], *b2 = (const W *)t.by[2];
        const double *dpg = fx ? c->scal : nullptr;
        const int nx = c->tune.cnzy_nxcd;
        const int sel = (t.Dz <= 4 ? 0 : 1) + (full ? 0 : 2);
        switch (sel * 16 + Dym) {
#define CHS_ZY(d) \
    case 0 * 16 + d: hipLaunchKernelGGL((k_cn_zy_march<W, d, ZYKB, ZYBT, 4, false>), dim3(nb), dim3(64, ZYKB), 0, c->stream, g, (const W *)t.gz, t.Dz, b0, b1, b2, rhs, c->q, c->qs, KZY, dpg, nx); break; \
    case 1 * 16 + d: hipLaunchKernelGGL((k_cn_zy_march<W, d, ZYKB, ZYBT, 8, false>), dim3(nb), dim3(64, ZYKB), 0, c->stream, g, (const W *)t.gz, t.Dz, b0, b1, b2, rhs, c->q, c->qs, KZY, dpg, nx); break; \
    case 2 * 16 + d: hipLaunchKernelGGL((k_cn_zy_march<W, d, ZYKB, ZYBT, 4, true>), dim3(nb), dim3(64, ZYKB), 0, c->stream, g, (const W *)t.gz, t.Dz, b0, b1, b2, rhs, c->q, c->qs, KZY, dpg, nx); break; \
    case 3 * 16 + d: hipLaunchKernelGGL((k_cn_zy_march<W, d, ZYKB, ZYBT, 8, true>), dim3(nb), dim3(64, ZYKB), 0, c->stream, g, (const W *)t.gz, t.Dz, b0, b1, b2, rhs, c->q, c->qs, KZY, dpg, nx); break;
            CHS_ZY(0) CHS_ZY(1) CHS_ZY(2) CHS_ZY(3) CHS_ZY(4) CHS_ZY(5) CHS_ZY(6) CHS_ZY(7) CHS_ZY(8)
#undef CHS_ZY
        }
    } else if (xz) {
        // rhs2 already holds the x- and z-solved increments
    } else if (zmarch) {
        EvScope ev(c, "cn_z");
        const dim3 gz((L + 63) / 64, (M + 3) / 4, 3 * ((N + KZ - 1) / KZ));
        cn_z_dispatch<W>(t.Dz, gz, blk, c->stream, g, (const W *)t.gz, rhs, rhs2, KZ, fx ? c->scal : nullptr);
    } else {
        for (int comp = 0; comp < 3; ++comp) {
            EvScope ev(c, "cn_z");
            hipLaunchKernelGGL(k_cn_z<W>, grd, blk, 0, c->stream, g, (const W *)t.gz, t.Dz, t.fullz, rhs, rhs2, comp);
        }
    }
    const int KY3 = c->tune.cny3_ky > 0 ? c->tune.cny3_ky : M;
    if (zy) {
        // the y-solves ran inside k_cn_zy_march
    } else if (c->tune.cny3_ky >= 0 && Dym <= 8 && KY3 > Dym && ysame) {
        EvScope ev(c, "cn_y");
        const dim3 gy((L + 63) / 64, (N + 3) / 4, 3 * ((M + KY3 - 1) / KY3));
        const W *b0 = (const W *)t.by[0], *b1 = (const W *)t.by[1], *b2 = (const W *)t.by[2];
        switch (Dym) {
#define CHS_Y3CASE(d) case d: hipLaunchKernelGGL((k_cn_y3_march<W, d>), gy, blk, 0, c->stream, g, b0, b1, b2, rhs2, c->q, c->qs, KY3); break;
            CHS_Y3CASE(0) CHS_Y3CASE(1) CHS_Y3CASE(2) CHS_Y3CASE(3) CHS_Y3CASE(4) CHS_Y3CASE(5) CHS_Y3CASE(6)
            CHS_Y3CASE(7) CHS_Y3CASE(8)
#undef CHS_Y3CASE
        }
    } else for (int comp = 0; comp < 3; ++comp) {
        EvScope ev(c, "cn_y");
        if (t.Dy[comp] <= 8) {
            const dim3 gy((L + 63) / 64, (N + 3) / 4, (M + KY - 1) / KY);
            cn_y_dispatch<W>(t.Dy[comp], gy, blk, c->stream, g, (const W *)t.by[comp], rhs2, c->q, c->qs, comp, KY);
        } else {
            hipLaunchKernelGGL(k_cn_y<W>, grd, blk, 0, c->stream, g, (const W *)t.by[comp], t.Dy[comp], rhs2,
                               comp == 0 ? rhs2 + cs : nullptr, c->q, c->qs, comp);
        }
    }
    // Poisson RHS from the prediction (c->qs); the correction recomputes the divergence
    poisson_launch<W>(c, c->dstep, c->talp[ns], nullptr, 0, c->qs, false);
    // FP64-residual refinement (mixed form): the divergence of the corrected velocity drives
    // another solve and correction; the CFL then comes from the final field
    const int nref = ns == 0 ? c->nref_projection : c->nref_step;
    const int KU = c->tune.upd_kz, UY = c->tune.upd_by;
    const bool u2 = c->tune.upd_w2 && L % 128 == 0;   // two points per lane
    const dim3 gu(u2 ? L / 128 : (L + 63) / 64, (M + UY - 1) / UY, (N + KU - 1) / KU), blku(64, UY);
    if (u2) {
        EvScope ev(c, "update");
        if (cfl && nref == 0)
            hipLaunchKernelGGL(k_update2<true>, gu, blku, 0, c->stream, g, s, c->phi, c->qs, c->q, c->pr, KU,
                               c->partial);
        else
            hipLaunchKernelGGL(k_update2<false>, gu, blku, 0, c->stream, g, s, c->phi, c->qs, c->q, c->pr, KU,
                               c->partial);
    } else {
        EvScope ev(c, "update");
        if (cfl && nref == 0)
            hipLaunchKernelGGL(k_update<true>, gu, blku, 0, c->stream, g, s, c->phi, c->qs, c->q, c->pr, KU,
                               c->partial, c->tune.upd_nxcd);
        else
            hipLaunchKernelGGL(k_update<false>, gu, blku, 0, c->stream, g, s, c->phi, c->qs, c->q, c->pr, KU,
                               c->partial, c->tune.upd_nxcd);
    }
    for (int r = 0; r < nref; ++r) {
        poisson_launch<W>(c, c->dstep, c->talp[ns], nullptr, 0, c->q, false);
        EvScope ev(c, "refine_update");
        hipLaunchKernelGGL(k_update_refine, grd, blk, 0, c->stream, g, s, c->phi, c->q, c->pr);
    }
    if (cfl) {
        EvScope ev(c, "cfl");
        int np = gu.x * gu.y * gu.z;
        if (nref > 0) {
            hipLaunchKernelGGL(k_cfl_div, grd, blk, 0, c->stream, g, c->q, c->partial, 0);
            np = grd.x * grd.y * grd.z;
        }
        hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(1024), 0, c->stream, c->partial, np, c->dstep + 2);
    }
    c->cfl_valid = cfl;
    HIPCHK(c, hipGetLastError());
    return 0;
}

// Poisson solve of a device RHS (cells) into a device phi (cells) with `nref` refinement
// sweeps: r = rhs - L phi in FP64, phi += solve(r).  resmax (optional) gets max |r| after
// the last sweep.
template <typename T>
int poisson_device_impl(chs_ctx *c, const double *rhs, double *phi, int nref, double *resmax)
{
    const Geo &g = c->geo;
    const dim3 blk(64, 4), grd((c->L + 63) / 64, (c->M + 3) / 4, c->N);
    const int npart = grd.x * grd.y * grd.z;
    const double *acph = c->cf_acph;
    poisson_launch<T>(c, nullptr, 1.0, rhs);
    hipLaunchKernelGGL(k_phi_accum, grd, blk, 0, c->stream, g, c->phi, phi, 0, (const double *)nullptr);
    for (int r = 0; r < nref; ++r) {
        hipLaunchKernelGGL(k_poisson_resid, grd, blk, 0, c->stream, g, c->d_amph, acph, c->d_apph, rhs, phi,
                           c->div, c->partial);
        poisson_launch<T>(c, nullptr, 1.0, c->div);
        // FP32 transforms: restore the exact gauge in FP64 on the last sweep.
        const double *shift = nullptr;
        if (sizeof(T) == 4 && r == nref - 1) {
            hipLaunchKernelGGL(k_gauge_mean, dim3(1), dim3(256), 0, c->stream, g, phi, c->phi, c->scal + 4);
            shift = c->scal + 4;
        }
        hipLaunchKernelGGL(k_phi_accum, grd, blk, 0, c->stream, g, c->phi, phi, 1, shift);
    }
    if (resmax) {
        hipLaunchKernelGGL(k_poisson_resid, grd, blk, 0, c->stream, g, c->d_amph, acph, c->d_apph, rhs, phi,
                           (double *)nullptr, c->partial);
        hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(256), 0, c->stream, c->partial, npart, c->scal + 3);
        HIPCHK(c, hipMemcpyAsync(resmax, c->scal + 3, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    HIPCHK(c, hipGetLastError());
    return 0;
}

template <typename T>
int poisson_impl(chs_ctx *c, const double *rhs, double *phi, int gauge)
{
    const int L = c->L, M = c->M, N = c->N;
    const size_t cs = (size_t)N * M * L, hs = (size_t)N * (M + 2) * L;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double *src = nullptr;
    HIPCHK(c, hipMalloc(&src, sizeof(double) * cs));
    hipError_t e = hipMemcpyAsync(src, rhs, sizeof(double) * cs, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        poisson_launch<T>(c, nullptr, 1.0, src, gauge);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    std::vector<double> h(hs);
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), c->phi, sizeof(double) * hs, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(src);
    if (e != hipSuccess) return fail(c, CHS_E_HIP, std::string("chs_poisson: ") + hipGetErrorString(e));
    for (int k = 0; k < N; ++k)
        for (int J = 1; J <= M; ++J)
            std::memcpy(phi + ((size_t)k * M + J - 1) * L, h.data() + ((size_t)k * (M + 2) + J) * L,
                        sizeof(double) * L);
    return 0;
}

}  // namespace

// =========================================================================================
// C-ABI
// =========================================================================================
extern "C" {

const char *chs_version(void) { return CHS_VERSION; }

const char *chs_last_error(const chs_ctx *ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

int chs_channel_setup(int L, int M, int N, double HX, double HZ, double STR2, double *scal, double *dyfi,
                      double *dyci, double *wff, double *wfb, double *amvr, double *acvr, double *apvr,
                      double *amph, double *acph, double *apph, double *ak1, double *ak3, double *tgam,
                      double *troh, double *talp)
{
    if (L < 4 || M < 4 || N < 4 || !(HX > 0) || !(HZ > 0)) return fail(nullptr, CHS_E_ARG, "bad mesh size");
    // CONSPARA.f90:44-73
    const double dx = HX / (double)L, dz = HZ / (double)N;
    const double dxi = 1.0 / dx, dzi = 1.0 / dz;
    // COORDJR.f90:44-55: YND(J) = tanh(STR2 (x2 - 1/2)) / tanh(STR2/2), x2 = (J-1)/(NND2-1)
    std::vector<double> ynd(M + 1), ycc(M);
    for (int J = 1; J <= M + 1; ++J) {
        const double x2 = (double)(J - 1) / (double)M;
        ynd[J - 1] = std::tanh(STR2 * (x2 - 0.5)) / std::tanh(STR2 * 0.5);
    }
    for (int J = 1; J <= M; ++J) ycc[J - 1] = (ynd[J - 1] + ynd[J]) * 0.5;
    // COORDJR.f90:121-135
    for (int J = 1; J <= M; ++J) dyfi[J - 1] = 1.0 / (ynd[J] - ynd[J - 1]);
    for (int J = 2; J <= M; ++J) dyci[J - 1] = 1.0 / ((ynd[J] - ynd[J - 2]) * 0.5);
    dyci[0] = 2.0 / (ynd[1] - ynd[0]);
    dyci[M] = 2.0 / (ynd[M] - ynd[M - 1]);
    // COORDJR.f90:425-442
    for (int J = 2; J <= M; ++J) {
        wff[J - 1] = (ynd[J - 1] - ycc[J - 2]) / (ycc[J - 1] - ycc[J - 2]);
        wfb[J - 1] = (ycc[J - 1] - ynd[J - 1]) / (ycc[J - 1] - ycc[J - 2]);
    }
    wfb[0] = 1.0; wff[0] = 0.0; wfb[M] = 0.0; wff[M] = 1.0;
    // COORDJR.f90:380-388
    double volm = 0.0;
    for (int k = 0; k < N; ++k)
        for (int J = 1; J <= M; ++J)
            for (int i = 0; i < L; ++i) volm = volm + 1.0 / dyfi[J - 1];
    // TDMA_COEF.f90:44-65
    for (int J = 1; J <= M; ++J) {
        amph[J - 1] = dyfi[J - 1] * dyci[J - 1];
        apph[J - 1] = dyfi[J - 1] * dyci[J];
        if (J == 1) amph[J - 1] = dyfi[J - 1] * dyci[J - 1] * 0.5;
        if (J + 1 == M + 1) apph[J - 1] = dyfi[J - 1] * dyci[J] * 0.5;
    }
    for (int J = 1; J <= M; ++J) acph[J - 1] = -(amph[J - 1] + apph[J - 1]);
    acph[0] = amph[0] + acph[0];
    amph[0] = 0.0;
    acph[M - 1] = apph[M - 1] + acph[M - 1];
    apph[M - 1] = 0.0;
    // TDMA_COEF.f90:67-104, A*VR(1:NND2, 1:3) column-major
    const int S = M + 1;
    for (int n = 0; n < 3 * S; ++n) amvr[n] = acvr[n] = apvr[n] = 0.0;
    auto DYFI = [&](int J) { return dyfi[J - 1]; };
    auto DYCI = [&](int J) { return dyci[J - 1]; };
    for (int idr = 1; idr <= 3; ++idr) {
        double *am = amvr + (idr - 1) * S - 1, *ac = acvr + (idr - 1) * S - 1, *ap = apvr + (idr - 1) * S - 1;
        if (idr != 2) {
            for (int J = 2; J <= M - 1; ++J) {
                am[J] = DYFI(J) * DYCI(J);
                ap[J] = DYFI(J) * DYCI(J + 1);
                ac[J] = -(am[J] + ap[J]);
            }
            double ucaj = 4.0 / (2.0 / DYCI(2) + 1.0 / DYFI(1));
            am[1] = 0.0;
            ap[1] = ucaj * DYCI(2);
            ac[1] = -ucaj * (DYCI(2) + 2.0 * DYFI(1));
            ucaj = 4.0 / (2.0 / DYCI(M) + 1.0 / DYFI(M));
            am[M] = ucaj * DYCI(M);
            ap[M] = 0.0;
            ac[M] = -ucaj * (2.0 * DYFI(M) + DYCI(M));
        } else {
            for (int J = 2; J <= M; ++J) {
                am[J] = DYFI(J - 1) * DYCI(J);
                ap[J] = DYFI(J) * DYCI(J);
                ac[J] = -(am[J] + ap[J]);
            }
            am[1] = ap[1] = 0.0;
            ac[1] = 1.0;
            am[M + 1] = ap[M + 1] = 0.0;
            ac[M + 1] = 1.0;
        }
    }
    // possion3d_FFT99.f90:171-199
    const double pi = 2.0 * std::asin(1.0);
    for (int K = 1; K <= N; ++K) {
        double an = (K <= N / 2) ? (double)(K - 1) * 2.0 * pi : -(double)(N - K + 1) * 2.0 * pi;
        ak3[K - 1] = 2.0 * (1.0 - std::cos(an / (double)N)) * (dzi * dzi);
    }
    for (int I = 1; I <= L; ++I) {
        double ap = (I <= L / 2) ? (double)(I - 1) * 2.0 * pi : -(double)(L - I + 1) * 2.0 * pi;
        ak1[I - 1] = 2.0 * (1.0 - std::cos(ap / (double)L)) * (dxi * dxi);
    }
    // CONS_RKCOEF.f90:21-51
    const double tg[4] = {1.0, 8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
    const double tr[4] = {0.0, 0.0, -17.0 / 60.0, -5.0 / 12.0};
    for (int n = 0; n < 4; ++n) { tgam[n] = tg[n]; troh[n] = tr[n]; talp[n] = tg[n] + tr[n]; }
    scal[0] = dxi; scal[1] = dzi; scal[2] = dxi * dxi; scal[3] = dzi * dzi; scal[4] = volm;
    return 0;
}

static int create_impl(const chs_coeffs *cf, int mode, hipStream_t shared, chs_ctx **out);

int chs_create(const chs_coeffs *cf, int mode, chs_ctx **out) { return create_impl(cf, mode, nullptr, out); }

static int create_impl(const chs_coeffs *cf, int mode, hipStream_t shared, chs_ctx **out)
{
    if (!cf || !out) return fail(nullptr, CHS_E_ARG, "null argument");
    *out = nullptr;
    if (mode != CHS_PLAIN_FP64 && mode != CHS_MIXED) return fail(nullptr, CHS_E_ARG, "bad mode");
    const int L = cf->L, M = cf->M, N = cf->N;
    chs::FftPlan px, pz;
    if (L < 4 || M < 4 || N < 4 || (L & 1) || (N & 1) || L > 2048 || N > 2048 || !fft_ok(L, px) ||
        !fft_ok(N, pz))
        return fail(nullptr, CHS_E_ARG, "L, N must be even, <= 2048, with factors 2,3,5 only; M >= 4");
    chs_ctx *c = new chs_ctx();
    c->L = L; c->M = M; c->N = N; c->H = L / 2 + 1; c->mode = mode;
    if (mode == CHS_MIXED) {   // FP32 stencils
        c->tune.cnz_kz = 8;
        c->tune.cny_ky = 32;
        // k_conv_rhs (where k_conv_rhs_lds does not apply): loads of a plane ahead of its
        // arithmetic and XCD-grouped tiles, 1.2 % faster per step at 256 x 192 x 256 in this form
        c->tune.conv_fence = 1;
        c->tune.nxcd = 8;
    }
    if (mode == CHS_MIXED) c->nref_projection = 2;
    c->px = px; c->pz = pz;
    c->dxi = cf->dxi; c->dzi = cf->dzi; c->dxqi = cf->dxqi; c->dzqi = cf->dzqi; c->volm = cf->volm;
    c->dyfi.assign(cf->dyfi, cf->dyfi + M);
    c->dyci.assign(M + 2, 0.0);
    c->wff.assign(M + 2, 0.0);
    c->wfb.assign(M + 2, 0.0);
    for (int J = 1; J <= M + 1; ++J) {
        c->dyci[J] = cf->dyci[J - 1];
        c->wff[J] = cf->wff[J - 1];
        c->wfb[J] = cf->wfb[J - 1];
    }
    c->amvr.assign(3 * (M + 2), 0.0);
    c->acvr.assign(3 * (M + 2), 0.0);
    c->apvr.assign(3 * (M + 2), 0.0);
    for (int comp = 0; comp < 3; ++comp)
        for (int J = 1; J <= M + 1; ++J) {
            c->amvr[comp * (M + 2) + J] = cf->amvr[comp * (M + 1) + J - 1];
            c->acvr[comp * (M + 2) + J] = cf->acvr[comp * (M + 1) + J - 1];
            c->apvr[comp * (M + 2) + J] = cf->apvr[comp * (M + 1) + J - 1];
        }
    c->amph.assign(cf->amph, cf->amph + M);
    c->acph.assign(cf->acph, cf->acph + M);
    c->apph.assign(cf->apph, cf->apph + M);
    c->ak1.assign(cf->ak1, cf->ak1 + L);
    c->ak3.assign(cf->ak3, cf->ak3 + N);
    for (int n = 0; n < 4; ++n) { c->tgam[n] = cf->tgam[n]; c->troh[n] = cf->troh[n]; c->talp[n] = cf->talp[n]; }

    auto bad = [&](int code) { chs_destroy(c); return code; };
#define CK(expr)                                                                        \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess) {                                                         \
            fail(nullptr, CHS_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
            return bad(CHS_E_HIP);                                                      \
        }                                                                               \
    } while (0)
    if (shared) {
        c->stream = shared;
        c->own_stream = false;
    } else {
        CK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    }
    // packed coefficients
    std::vector<double> pack;
    auto put = [&](const std::vector<double> &v) { size_t o = pack.size(); pack.insert(pack.end(), v.begin(), v.end()); return o; };
    size_t o_dyfi = put(c->dyfi), o_dyci = put(c->dyci), o_wff = put(c->wff), o_wfb = put(c->wfb);
    size_t o_am = put(c->amvr), o_ac = put(c->acvr), o_ap = put(c->apvr);
    size_t o_amph = put(c->amph), o_apph = put(c->apph), o_acph = put(c->acph);
    CK(hipMalloc(&c->cf, sizeof(double) * pack.size()));
    CK(hipMemcpyAsync(c->cf, pack.data(), sizeof(double) * pack.size(), hipMemcpyHostToDevice, c->stream));
    CK(hipStreamSynchronize(c->stream));
    Geo &g = c->geo;
    g.L = L; g.M = M; g.N = N; g.H = c->H;
    g.dxi = c->dxi; g.dzi = c->dzi; g.dxqi = c->dxqi; g.dzqi = c->dzqi;
    g.dyfi = c->cf + o_dyfi; g.dyci = c->cf + o_dyci; g.wff = c->cf + o_wff; g.wfb = c->cf + o_wfb;
    g.amvr = c->cf + o_am; g.acvr = c->cf + o_ac; g.apvr = c->cf + o_ap;
    c->d_amph = c->cf + o_amph; c->d_apph = c->cf + o_apph; c->cf_acph = c->cf + o_acph;

    const size_t hs = (size_t)N * (M + 2) * L, cs = (size_t)N * M * L;
    const size_t ws = wsize(c);
    CK(hipMalloc(&c->q, sizeof(double) * 3 * hs));
    CK(hipMalloc(&c->pr, sizeof(double) * hs));
    CK(hipMalloc(&c->phi, sizeof(double) * hs));
    CK(hipMalloc(&c->div, sizeof(double) * cs));
    CK(hipMalloc(&c->h0, ws * 3 * cs));
    CK(hipMalloc(&c->rhs, ws * 3 * cs));
    CK(hipMalloc(&c->rhs2, ws * 3 * cs));
    CK(hipMalloc(&c->spec, 2 * ws * (size_t)c->H * M * N));
    CK(hipMalloc(&c->pp, sizeof(double) * (size_t)c->H * M * N));
    // >= every k_conv_rhs grid chs_set_tuning allows (conv_by >= 1, conv_kc >= 1) and k_cfl_div's
    c->npartial = ((L + 63) / 64) * M * N;
    CK(hipMalloc(&c->partial, sizeof(double) * c->npartial));
    CK(hipMalloc(&c->scal, sizeof(double) * 16));
    c->dstep = c->scal + 8;
    CK(hipMalloc(&c->qs, sizeof(double) * 3 * hs));
    CK(hipMemsetAsync(c->qs, 0, sizeof(double) * 3 * hs, c->stream));
    CK(hipMemsetAsync(c->q, 0, sizeof(double) * 3 * hs, c->stream));
    CK(hipMemsetAsync(c->pr, 0, sizeof(double) * hs, c->stream));
    CK(hipMemsetAsync(c->phi, 0, sizeof(double) * hs, c->stream));
    CK(hipMemsetAsync(c->h0, 0, ws * 3 * cs, c->stream));
    CK(hipMemsetAsync(c->rhs, 0, ws * 3 * cs, c->stream));
    CK(hipMemsetAsync(c->rhs2, 0, ws * 3 * cs, c->stream));
    CK(hipMemsetAsync(c->scal, 0, sizeof(double) * 16, c->stream));
    // Poisson pivots (TDMAIJJ_nonCYC.f90:84-106 on TDMA_PHI_COEF, possion3d_FFT99.f90:213-231)
    {
        const int H = c->H;
        std::vector<double> pp((size_t)H * M * N);
        std::vector<double> hh(M);
        for (int kz = 0; kz < N; ++kz)
            for (int kx = 0; kx < H; ++kx) {
                double hprev = 0.0;
                for (int J = 1; J <= M; ++J) {
                    double a = c->amph[J - 1], cc = c->apph[J - 1];
                    double b = c->acph[J - 1] + (-c->ak1[kx] / 1.0 - c->ak3[kz]);
                    if (kx == 0 && kz == 0 && J == 1) { a = 0.0; cc = 0.0; b = c->acph[0] + (-c->ak1[0] / 1.0 - c->ak3[0]); }
                    double p = (J == 1) ? 1.0 / b : 1.0 / (b - hprev * a);
                    hprev = cc * p;
                    pp[((size_t)(J - 1) * H + kx) * N + kz] = p;
                }
            }
        CK(hipMemcpyAsync(c->pp, pp.data(), sizeof(double) * pp.size(), hipMemcpyHostToDevice, c->stream));
        CK(hipStreamSynchronize(c->stream));
        // (0,0) mode as TRID0 eliminates it (possion3d_fishpack.f90:310-340): unmodified
        // coefficients, rows 1..M-1; row M is the singular one whose unknown is pinned to 0
        std::vector<double> p00(M, 0.0);
        double hprev = 0.0;
        for (int J = 1; J < M; ++J) {
            const double a = c->amph[J - 1], cc = c->apph[J - 1], b = c->acph[J - 1];
            const double p = (J == 1) ? 1.0 / b : 1.0 / (b - hprev * a);
            hprev = cc * p;
            p00[J - 1] = p;
        }
        CK(hipMalloc(&c->pp00, sizeof(double) * M));
        CK(hipMemcpyAsync(c->pp00, p00.data(), sizeof(double) * M, hipMemcpyHostToDevice, c->stream));
        CK(hipStreamSynchronize(c->stream));
        if (mode == CHS_MIXED) {
            std::vector<float> p32(pp.begin(), pp.end());
            CK(hipMalloc(&c->pp32, sizeof(float) * p32.size()));
            CK(hipMemcpyAsync(c->pp32, p32.data(), sizeof(float) * p32.size(), hipMemcpyHostToDevice, c->stream));
            CK(hipStreamSynchronize(c->stream));
        }
    }
    int e = (mode == CHS_MIXED) ? upload_twiddles<float>(c, L, &c->twx) : upload_twiddles<double>(c, L, &c->twx);
    if (!e) e = (mode == CHS_MIXED) ? upload_twiddles<float>(c, N, &c->twz) : upload_twiddles<double>(c, N, &c->twz);
    // register FFTs: x needs a plan for L and N divisible by the plane pairs of one wavefront
    int rad[8];
    if (!e && chs::rplan_radices(L, rad) && N % (2 * chs::rplan_lpw(L)) == 0)
        e = (mode == CHS_MIXED) ? upload_rtw<float>(c, L, &c->rtwx) : upload_rtw<double>(c, L, &c->rtwx);
    if (!e && chs::rplan_radices(N, rad))
        e = (mode == CHS_MIXED) ? upload_rtw<float>(c, N, &c->rtwz) : upload_rtw<double>(c, N, &c->rtwz);
    if (e) return bad(e);
#undef CK
    *out = c;
    return 0;
}

void chs_destroy(chs_ctx *c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    void *ptrs[] = {c->cf, c->q, c->qs, c->pr, c->phi, c->div, c->h0, c->rhs, c->rhs2, c->spec, c->pp,
                    c->pp32, c->pp00, c->partial, c->scal, c->twx, c->twz, c->rtwx, c->rtwz, c->cn_stamp};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (auto &t : c->cn) {
        if (t.gx) (void)hipFree(t.gx);
        if (t.gz) (void)hipFree(t.gz);
        for (auto *b : t.by)
            if (b) (void)hipFree(b);
    }
    for (auto &u : c->ev_used) { (void)hipEventDestroy(u.second.first); (void)hipEventDestroy(u.second.second); }
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->stream && c->own_stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int chs_upload_fields(chs_ctx *c, const double *q, const double *pr, const double *convh0)
{
    if (!c || !q || !pr) return fail(c, CHS_E_ARG, "null argument");
    const size_t hs = (size_t)c->N * (c->M + 2) * c->L, cs = (size_t)c->N * c->M * c->L;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpyAsync(c->q, q, sizeof(double) * 3 * hs, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->pr, pr, sizeof(double) * hs, hipMemcpyHostToDevice, c->stream));
    if (convh0) {
        if (c->mode == CHS_MIXED) {
            std::vector<float> h(convh0, convh0 + 3 * cs);
            HIPCHK(c, hipMemcpyAsync(c->h0, h.data(), sizeof(float) * 3 * cs, hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
        } else {
            HIPCHK(c, hipMemcpyAsync(c->h0, convh0, sizeof(double) * 3 * cs, hipMemcpyHostToDevice, c->stream));
        }
    } else {
        HIPCHK(c, hipMemsetAsync(c->h0, 0, wsize(c) * 3 * cs, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->have_fields = true;
    c->cfl_valid = false;
    return 0;
}

int chs_download_fields(chs_ctx *c, double *q, double *pr, double *convh0)
{
    if (!c) return fail(c, CHS_E_ARG, "null ctx");
    const size_t hs = (size_t)c->N * (c->M + 2) * c->L, cs = (size_t)c->N * c->M * c->L;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (q) HIPCHK(c, hipMemcpyAsync(q, c->q, sizeof(double) * 3 * hs, hipMemcpyDeviceToHost, c->stream));
    if (pr) HIPCHK(c, hipMemcpyAsync(pr, c->pr, sizeof(double) * hs, hipMemcpyDeviceToHost, c->stream));
    if (convh0) {
        if (c->mode == CHS_MIXED) {
            std::vector<float> h(3 * cs);
            HIPCHK(c, hipMemcpyAsync(h.data(), c->h0, sizeof(float) * 3 * cs, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            for (size_t n = 0; n < 3 * cs; ++n) convh0[n] = h[n];
        } else {
            HIPCHK(c, hipMemcpyAsync(convh0, c->h0, sizeof(double) * 3 * cs, hipMemcpyDeviceToHost, c->stream));
        }
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int chs_substep_tg(chs_ctx *c, int ns, double dt, double cvisc)
{
    if (!c) return fail(c, CHS_E_ARG, "null ctx");
    if (ns < 0 || ns > 3) return fail(c, CHS_E_ARG, "NS must be 0..3");
    if (!(dt > 0) || !(cvisc > 0)) return fail(c, CHS_E_ARG, "dt and cvisc must be positive");
    if (!c->have_fields) return fail(c, CHS_E_STATE, "no fields uploaded");
    return c->mode == CHS_MIXED ? substep_impl<float>(c, ns, dt, cvisc) : substep_impl<double>(c, ns, dt, cvisc);
}

int chs_step_tg(chs_ctx *c, double dt, double cvisc)
{
    if (!c) return fail(c, CHS_E_ARG, "null ctx");
    if (!(dt > 0) || !(cvisc > 0)) return fail(c, CHS_E_ARG, "dt and cvisc must be positive");
    if (!c->have_fields) return fail(c, CHS_E_STATE, "no fields uploaded");
    auto run = [&](int ns, bool host) {
        return c->mode == CHS_MIXED ? substep_impl<float>(c, ns, dt, cvisc, host, ns == 3)
                                    : substep_impl<double>(c, ns, dt, cvisc, host, ns == 3);
    };
    if (!c->use_graph || c->prof) {
        for (int ns = 1; ns <= 3; ++ns)
            if (int e = run(ns, true)) return e;
        return 0;
    }
    // widths and tables first (no-ops when current): a change of widths changes the kernels,
    // which a captured graph cannot follow
    for (int ns = 1; ns <= 3; ++ns)
        if (int e = build_cn(c, ns, dt, cvisc)) return e;
    if (!c->gexec || c->gdt != dt || c->gcv != cvisc || c->gtune != c->tune_gen || c->gcn != c->cn_gen ||
        c->gkind != 0) {
        if (c->gexec) {
            HIPCHK(c, hipGraphExecDestroy(c->gexec));
            c->gexec = nullptr;
        }
        hipGraph_t graph = nullptr;
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        // the graph sets DT, CVISC and (where their stamps differ) the tables itself, so a
        // replay does not depend on what other calls left on the device
        c->dv_host = false;
        set_step(c, dt, cvisc);
        cn_launch(c, 1, 3);
        int e = 0;
        for (int ns = 1; ns <= 3 && !e; ++ns) e = run(ns, false);
        hipError_t ce = hipStreamEndCapture(c->stream, &graph);
        if (e) {
            if (graph) (void)hipGraphDestroy(graph);
            return e;
        }
        if (ce != hipSuccess) return fail(c, CHS_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
        ce = hipGraphInstantiate(&c->gexec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ce != hipSuccess) return fail(c, CHS_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ce));
        c->gdt = dt;
        c->gcv = cvisc;
        c->gtune = c->tune_gen;
        c->gcn = c->cn_gen;
        c->gkind = 0;
    }
    HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
    c->dv_host = true;
    c->dv_dt = dt;
    c->dv_cv = cvisc;
    c->cfl_valid = true;
    return 0;
}

int chs_run_tg(chs_ctx *c, int nsteps, double dt0, double cflgv, double dtmin, double cvisc)
{
    if (!c) return fail(c, CHS_E_ARG, "null ctx");
    if (nsteps < 0 || !(dt0 > 0) || !(cflgv > 0) || !(dtmin >= 0) || !(cvisc > 0))
        return fail(c, CHS_E_ARG, "chs_run_tg: nsteps >= 0, dt0 > 0, cflgv > 0, dtmin >= 0, cvisc > 0");
    if (!c->have_fields) return fail(c, CHS_E_STATE, "no fields uploaded");
    if (!c->dstate) return fail(c, CHS_E_STATE, "chs_run_tg: set the step state first (chs_set_step_state)");
    if (int e = cn_alloc(c)) return e;
    // widths for the largest DT the rule allows (DT0): every smaller DT needs no wider band
    for (int ns = 1; ns <= 3; ++ns) {
        CnTab &t = c->cn[ns];
        CnTab w = t;
        cn_widths(c, ns, dt0, cvisc, w);
        const bool changed = w.Dx != t.Dx || w.Dz != t.Dz || w.fullx != t.fullx || w.fullz != t.fullz ||
                             w.Dy[0] != t.Dy[0] || w.Dy[1] != t.Dy[1] || w.Dy[2] != t.Dy[2];
        if (changed) {
            ++c->cn_gen;
            t.Dx = w.Dx; t.Dz = w.Dz; t.fullx = w.fullx; t.fullz = w.fullz;
            for (int k = 0; k < 3; ++k) t.Dy[k] = w.Dy[k];
            hipLaunchKernelGGL(k_cn_stamp, dim3(1), dim3(64), 0, c->stream, nullptr, c->cn_stamp, ns, 1);
        }
        t.valid = false;   // the device, not the host, chooses the DT these tables hold
    }
    if (!c->cfl_valid) {   // CFLMM of the current field for the first step's rule
        const dim3 blk(64, 4), grd((c->L + 63) / 64, (c->M + 3) / 4, c->N);
        hipLaunchKernelGGL(k_cfl_div, grd, blk, 0, c->stream, c->geo, c->q, c->partial, 0);
        hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(1024), 0, c->stream, c->partial,
                           (int)(grd.x * grd.y * grd.z), c->dstep + 2);
        c->cfl_valid = true;
    }
    c->dv_host = false;
    auto step = [&]() -> int {
        hipLaunchKernelGGL(k_dt_rule, dim3(1), dim3(64), 0, c->stream, c->dstep, dt0, cflgv, dtmin, cvisc);
        cn_launch(c, 1, 3);
        for (int ns = 1; ns <= 3; ++ns) {
            const int e = c->mode == CHS_MIXED ? substep_impl<float>(c, ns, 0.0, cvisc, false, ns == 3)
                                               : substep_impl<double>(c, ns, 0.0, cvisc, false, ns == 3);
            if (e) return e;
        }
        return 0;
    };
    if (!c->use_graph || c->prof) {
        for (int n = 0; n < nsteps; ++n)
            if (int e = step()) return e;
        HIPCHK(c, hipGetLastError());
        return 0;
    }
    const double key[4] = {dt0, cflgv, dtmin, cvisc};
    if (!c->gexec || c->gkind != 1 || c->gtune != c->tune_gen || c->gcn != c->cn_gen ||
        std::memcmp(key, c->grule, sizeof key) != 0) {
        if (c->gexec) {
            HIPCHK(c, hipGraphExecDestroy(c->gexec));
            c->gexec = nullptr;
        }
        hipGraph_t graph = nullptr;
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int e = step();
        hipError_t ce = hipStreamEndCapture(c->stream, &graph);
        if (e) {
            if (graph) (void)hipGraphDestroy(graph);
            return e;
        }
        if (ce != hipSuccess) return fail(c, CHS_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
        ce = hipGraphInstantiate(&c->gexec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ce != hipSuccess) return fail(c, CHS_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ce));
        std::memcpy(c->grule, key, sizeof key);
        c->gtune = c->tune_gen;
        c->gcn = c->cn_gen;
        c->gkind = 1;
    }
    for (int n = 0; n < nsteps; ++n) HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
    return 0;
}

int chs_set_step_state(chs_ctx *c, double dt, double phytime)
{
    if (!c || !(dt > 0)) return fail(c, CHS_E_ARG, "chs_set_step_state: dt must be positive");
    const double v[3] = {dt, phytime, 0.0};
    // DT, PhyTIME and a cleared stop flag; CVISC stays (the rule writes it each step)
    HIPCHK(c, hipMemcpyAsync(c->dstep + 0, &v[0], sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->dstep + 3, &v[1], sizeof(double) * 2, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->dv_host = false;
    c->dstate = true;
    return 0;
}

int chs_get_step_state(chs_ctx *c, double *dt, double *phytime, double *cflmm, int *stopped)
{
    if (!c) return fail(c, CHS_E_ARG, "null ctx");
    double v[5];
    HIPCHK(c, hipMemcpyAsync(v, c->dstep, sizeof v, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (dt) *dt = v[0];
    if (phytime) *phytime = v[3];
    if (cflmm) *cflmm = c->cfl_valid ? v[2] : std::nan("");
    if (stopped) *stopped = v[4] != 0.0;
    return 0;
}

int chs_set_refinement(chs_ctx *c, int nref_step, int nref_projection)
{
    if (!c || nref_step < 0 || nref_step > 4 || nref_projection < 0 || nref_projection > 4)
        return fail(c, CHS_E_ARG, "refinement sweeps must be 0..4");
    c->nref_step = nref_step;
    c->nref_projection = nref_projection;
    ++c->tune_gen;   // captured graphs hold the old sweep count
    return 0;
}

void *chs_dev_malloc(size_t bytes)
{
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    return p;
}

int chs_dev_free(void *p)
{
    if (p && hipFree(p) != hipSuccess) return fail(nullptr, CHS_E_HIP, "hipFree failed");
    return 0;
}

int chs_memcpy(void *dst, const void *src, size_t bytes, int to_device)
{
    hipError_t e = hipMemcpy(dst, src, bytes, to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail(nullptr, CHS_E_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
    return 0;
}

int chs_set_tuning(chs_ctx *c, const char *key, int value)
{
    if (!c || !key) return fail(c, CHS_E_ARG, "null argument");
    const std::string k(key);
    const int L = c->L, M = c->M, N = c->N;
    auto bad = [&](const char *why) { return fail(c, CHS_E_ARG, "chs_set_tuning " + k + ": " + why); };
    if (k == "conv_by") {
        if (value != 1 && value != 2 && value != 4 && value != 8 && value != 16) return bad("must be 1, 2, 4, 8 or 16");
        if (c->tune.conv_bx * value > 16) return bad("conv_bx * conv_by must be <= 16");
        c->tune.conv_by = value;
    } else if (k == "conv_fence") {
        if (value != 0 && value != 1) return bad("must be 0 or 1");
        c->tune.conv_fence = value;
    } else if (k == "conv_bx") {
        if (value != 1 && value != 4) return bad("must be 1 or 4");
        if (value * c->tune.conv_by > 16) return bad("conv_bx * conv_by must be <= 16");
        c->tune.conv_bx = value;
    } else if (k == "conv_kc") {
        if (value < 1 || value > N) return bad("must be 1..N");
        c->tune.conv_kc = value;
    } else if (k == "conv_fx") {
        if (value < 0 || value > 2) return bad("conv_fx must be 0, 1 or 2 (rows per block)");
        c->tune.conv_fx = value;
    } else if (k == "cnx_nlx") {
        if (value < 0 || value > M * N || (size_t)value * L * sizeof(double) > 65536) return bad("lines must fit 64 KiB of LDS");
        c->tune.cnx_nlx = value;
    } else if (k == "cnz_kz") {
        if (value < 1 || value > N) return bad("must be 1..N");
        c->tune.cnz_kz = value;
    } else if (k == "cny_ky") {
        if (value < 1 || value > M) return bad("must be 1..M");
        c->tune.cny_ky = value;
    } else if (k == "fftx_tk") {
        if (value != 0 && (value < 2 || (value & 1) || N % value || (value / 2) * L > chs::kMaxPts * 256))
            return bad("must be even, divide N, and (value/2)*L <= 2048");
        c->tune.fftx_tk = value;
    } else if (k == "cn_xz") {
        if (value != 0 && value != 1) return bad("must be 0 or 1");
        c->tune.cn_xz = value;
    } else if (k == "cnxz_by") {
        if (value != 1 && value != 2 && value != 4) return bad("must be 1, 2 or 4");
        c->tune.cnxz_by = value;
    } else if (k == "fft_reg") {
        if (value != 0 && value != 1) return bad("must be 0 or 1");
        c->tune.fft_reg = value;
    } else if (k == "nxcd") {
        if (value < 1 || value > 64) return bad("must be 1..64");
        c->tune.nxcd = value;
    } else if (k == "ytdma_bs") {
        if (value != 64 && value != 128 && value != 256) return bad("must be 64, 128 or 256");
        c->tune.ytdma_bs = value;
    } else if (k == "upd_by") {
        if (value != 1 && value != 2 && value != 4 && value != 8 && value != 16) return bad("must be 1, 2, 4, 8 or 16");
        c->tune.upd_by = value;
    } else if (k == "upd_nxcd") {
        if (value < 1 || value > 64) return bad("must be 1..64");
        c->tune.upd_nxcd = value;
    } else if (k == "upd_kz") {
        if (value < 1 || value > N) return bad("must be 1..N");
        c->tune.upd_kz = value;
    } else if (k == "conv_lds") {
        if (value != 0 && value != 1 && value != 2 && value != 4) return bad("must be 0, 1, 2 or 4");
        c->tune.conv_lds = value;
    } else if (k == "divg_nxcd") {
        if (value < 1 || value > 64) return bad("must be 1..64");
        c->tune.divg_nxcd = value;
    } else if (k == "upd_w2") {
        if (value < 0 || value > 1) return bad("must be 0 or 1");
        c->tune.upd_w2 = value;
    } else if (k == "cnzy") {
        if (value < 0 || value > 2) return bad("must be 0, 1 or 2");
        c->tune.cnzy = value;
    } else if (k == "cnzy_ky") {
        if (value < 0 || value > M) return bad("must be 0..M");
        c->tune.cnzy_ky = value;
    } else if (k == "cnzy_nxcd") {
        if (value < 1 || value > 64) return bad("must be 1..64");
        c->tune.cnzy_nxcd = value;
    } else if (k == "cny3_ky") {
        if (value < -1 || value > M) return bad("must be -1..M");
        c->tune.cny3_ky = value;
    } else if (k == "conv_dma") {
        if (value != 0 && value != 2 && value != 4) return bad("must be 0, 2 or 4");
        c->tune.conv_dma = value;
    } else if (k == "ytdma_c") {
        if (value != -1 && value != 0 && value != 4 && value != 8 && value != 16) return bad("must be -1, 0, 4, 8 or 16");
        c->tune.ytdma_c = value;
    } else if (k == "ytdma_rb") {
        if (value != 8 && value != 16) return bad("must be 8 or 16");
        c->tune.ytdma_rb = value;
    } else if (k == "fftz_nl") {
        if (value < 0 || value * N > chs::kMaxPts * 256) return bad("value*N must be <= 2048");
        c->tune.fftz_nl = value;
    } else {
        return bad("unknown key");
    }
    ++c->tune_gen;
    c->gdt = 0.0;   // force re-capture of a step graph
    return 0;
}

int chs_set_graph(chs_ctx *c, int enable)
{
    if (!c) return fail(c, CHS_E_ARG, "null ctx");
    c->use_graph = enable != 0;
    return 0;
}

int chs_poisson(chs_ctx *c, const double *rhs, double *phi)
{
    if (!c || !rhs || !phi) return fail(c, CHS_E_ARG, "null argument");
    return c->mode == CHS_MIXED ? poisson_impl<float>(c, rhs, phi, 0) : poisson_impl<double>(c, rhs, phi, 0);
}

int chs_poisson_fishpack(chs_ctx *c, const double *rhs, double *phi)
{
    if (!c || !rhs || !phi) return fail(c, CHS_E_ARG, "null argument");
    return c->mode == CHS_MIXED ? poisson_impl<float>(c, rhs, phi, 1) : poisson_impl<double>(c, rhs, phi, 1);
}

int chs_poisson_device(chs_ctx *c, const double *rhs, double *phi, int nref, double *resmax)
{
    if (!c || !rhs || !phi || nref < 0 || nref > 8) return fail(c, CHS_E_ARG, "bad argument");
    return c->mode == CHS_MIXED ? poisson_device_impl<float>(c, rhs, phi, nref, resmax)
                                : poisson_device_impl<double>(c, rhs, phi, nref, resmax);
}

int chs_cfl_tg(chs_ctx *c, double dt, double *cflmm)
{
    if (!c || !cflmm) return fail(c, CHS_E_ARG, "null argument");
    if (!c->cfl_valid) {   // a step forms CFLMM_tg of its new field; otherwise compute it here
        const Geo &g = c->geo;
        dim3 blk(64, 4), grd((c->L + 63) / 64, (c->M + 3) / 4, c->N);
        hipLaunchKernelGGL(k_cfl_div, grd, blk, 0, c->stream, g, c->q, c->partial, 0);
        hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(1024), 0, c->stream, c->partial,
                           (int)(grd.x * grd.y * grd.z), c->dstep + 2);
        c->cfl_valid = true;
    }
    double v = 0.0;
    HIPCHK(c, hipMemcpyAsync(&v, c->dstep + 2, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *cflmm = v * dt;
    if (!std::isfinite(v)) return fail(c, CHS_E_NONFINITE, "chs_cfl_tg: non-finite velocity");
    return 0;
}

int chs_max_div(chs_ctx *c, double *maxdiv)
{
    if (!c || !maxdiv) return fail(c, CHS_E_ARG, "null argument");
    const Geo &g = c->geo;
    dim3 blk(64, 4), grd((c->L + 63) / 64, (c->M + 3) / 4, c->N);
    hipLaunchKernelGGL(k_cfl_div, grd, blk, 0, c->stream, g, c->q, c->partial, 1);
    hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(1024), 0, c->stream, c->partial, (int)(grd.x * grd.y * grd.z),
                       c->scal + 2);
    double v = 0.0;
    HIPCHK(c, hipMemcpyAsync(&v, c->scal + 2, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *maxdiv = v;
    if (!std::isfinite(v)) return fail(c, CHS_E_NONFINITE, "chs_max_div: non-finite divergence");
    return 0;
}

int chs_sync(chs_ctx *c)
{
    if (!c) return fail(c, CHS_E_ARG, "null ctx");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int chs_device_fields(chs_ctx *c, double **q, double **pr)
{
    if (!c) return fail(c, CHS_E_ARG, "null ctx");
    if (q) *q = c->q;
    if (pr) *pr = c->pr;
    c->cfl_valid = false;   // the caller may write the fields
    return 0;
}

int chs_profile_reset(chs_ctx *c, const char *only)
{
    if (!c) return fail(c, CHS_E_ARG, "null ctx");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (auto &u : c->ev_used) { (void)hipEventDestroy(u.second.first); (void)hipEventDestroy(u.second.second); }
    c->ev_used.clear();
    c->prof = true;
    c->prof_only = only ? only : "";
    return 0;
}

int chs_profile_get(chs_ctx *c, const char *kernel, double *avg_ms, int *count)
{
    if (!c || !kernel || !avg_ms || !count) return fail(c, CHS_E_ARG, "null argument");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double tot = 0.0;
    int n = 0;
    for (auto &u : c->ev_used)
        if (u.first == kernel) {
            float ms = 0.f;
            HIPCHK(c, hipEventElapsedTime(&ms, u.second.first, u.second.second));
            tot += ms;
            ++n;
        }
    *avg_ms = n ? tot / n : 0.0;
    *count = n;
    return 0;
}

// ---- Ensemble of independent channels (one hipGraph per RK3 step of all channels) ----

struct chs_ens {
    std::vector<chs_ctx *> ch;
    hipStream_t stream = nullptr;
    // batched: one launch per kernel kind and stage over all channels (k_multi); otherwise
    // (the default: faster on the uneven workload, see chs_ens_set_batched) each channel's own
    // launches on the lanes below
    bool batched = false;
    bool gbatched = false;          // the current graph is the batched one
    std::vector<void *> mbufs;      // device item tables of the batched graph
    // Capture lanes: channel i's launches go to lane i % lanes.size(), so the step graph holds
    // independent chains that the GPU runs side by side (small channels alone fill little of it).
    std::vector<hipStream_t> lanes;
    std::vector<hipEvent_t> ev;   // ev[0] fork, ev[1 + s] join of lane s
    hipGraphExec_t gexec = nullptr;
    std::vector<double> gdt, gcv;
    std::vector<int> gtune, gcn;
    std::string err;
};

int chs_ens_create(int n, const chs_coeffs *const *coeffs, int mode, chs_ens **out)
{
    if (n < 1 || !coeffs || !out) return fail(nullptr, CHS_E_ARG, "chs_ens_create: bad argument");
    *out = nullptr;
    chs_ens *e = new chs_ens();
    hipError_t he = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (he != hipSuccess) {
        delete e;
        return fail(nullptr, CHS_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(he));
    }
    for (int i = 0; i < n; ++i) {
        chs_ctx *c = nullptr;
        int rc = create_impl(coeffs[i], mode, e->stream, &c);
        if (rc) {
            chs_ens_destroy(e);
            return rc;
        }
        e->ch.push_back(c);
    }
    const int nl = std::min(n, 8);
    e->ev.assign(nl + 1, nullptr);
    for (int s = 0; s < nl && he == hipSuccess; ++s) {
        hipStream_t st = nullptr;
        he = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (he == hipSuccess) e->lanes.push_back(st);
    }
    for (int s = 0; s <= nl && he == hipSuccess; ++s) he = hipEventCreateWithFlags(&e->ev[s], hipEventDisableTiming);
    if (he != hipSuccess) {
        chs_ens_destroy(e);
        return fail(nullptr, CHS_E_HIP, std::string("chs_ens_create: ") + hipGetErrorString(he));
    }
    *out = e;
    return 0;
}

void chs_ens_destroy(chs_ens *e)
{
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    for (void *b : e->mbufs) (void)hipFree(b);
    for (auto *c : e->ch) chs_destroy(c);
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    for (auto st : e->lanes) (void)hipStreamDestroy(st);
    for (auto ev : e->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

int chs_ens_channel(chs_ens *e, int i, chs_ctx **ctx)
{
    if (!e || i < 0 || i >= (int)e->ch.size() || !ctx) return fail(nullptr, CHS_E_ARG, "chs_ens_channel: bad index");
    *ctx = e->ch[i];
    return 0;
}

}  // extern "C"

namespace {

// One k_multi launch: the channels' items and first blocks in device memory (owned by the
// ensemble), the block shape and the dynamic LDS size.
template <class K> struct MLaunch {
    const MItem<typename K::Args> *it = nullptr;
    const int *bstart = nullptr;
    int n = 0, nblk = 0;
    dim3 blk;
    size_t sh = 0;
    void launch(hipStream_t st) const
    {
        hipLaunchKernelGGL(k_multi<K>, dim3(nblk), blk, sh, st, it, bstart, n);
    }
};

template <class K>
int mlaunch_make(chs_ens *e, const std::vector<typename K::Args> &args, const std::vector<dim3> &grids, dim3 blk,
                 size_t sh, MLaunch<K> &ml)
{
    using It = MItem<typename K::Args>;
    const int n = (int)args.size();
    std::vector<It> items(n);
    std::vector<int> first(n);
    long long nb = 0;
    for (int i = 0; i < n; ++i) {
        items[i].a = args[i];
        items[i].gx = grids[i].x;
        items[i].gy = grids[i].y;
        items[i].gz = grids[i].z;
        first[i] = (int)nb;
        nb += (long long)grids[i].x * grids[i].y * grids[i].z;
    }
    if (nb < 1 || nb > 0x7fffffffLL) return fail(nullptr, CHS_E_ARG, "ensemble launch: block count out of range");
    const size_t bi = (sizeof(It) * n + 255) / 256 * 256;
    void *d = nullptr;
    hipError_t he = hipMalloc(&d, bi + sizeof(int) * n);
    if (he == hipSuccess) e->mbufs.push_back(d);
    if (he == hipSuccess) he = hipMemcpy(d, items.data(), sizeof(It) * n, hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemcpy((char *)d + bi, first.data(), sizeof(int) * n, hipMemcpyHostToDevice);
    if (he != hipSuccess) return fail(nullptr, CHS_E_HIP, std::string("ensemble launch tables: ") + hipGetErrorString(he));
    ml.it = (const It *)d;
    ml.bstart = (const int *)((char *)d + bi);
    ml.n = n;
    ml.nblk = (int)nb;
    ml.blk = blk;
    ml.sh = sh;
    return 0;
}

// Whether every channel's step fits the batched launches: banded z and y solves (one y width),
// chunked y-Thomas (M <= 256), no refinement sweeps inside a step, the fused y-solve enabled.
bool ens_batchable(const chs_ens *e)
{
    for (const chs_ctx *c : e->ch)
        for (int ns = 1; ns <= 3; ++ns) {
            const CnTab &t = c->cn[ns];
            if (t.fullz || t.Dz > 8 || t.Dy[0] > 8 || t.Dy[1] != t.Dy[0] || t.Dy[2] != t.Dy[0]) return false;
            if (c->M > 256 || c->nref_step != 0 || c->tune.cny3_ky < 0 ||
                ytdma_chunk_rows(c->tune.ytdma_c, c->M, c->H * c->N) == 0)
                return false;
            if (c->tune.cny3_ky > 0 && c->tune.cny3_ky <= t.Dy[0]) return false;
        }
    return true;
}

// The batched step graph's launches, in a channel's own order (substep_impl with the x-solve in
// its own kernel, the LDS transforms, 64 x 4 stencil blocks): per stage convection, mass-flux
// sum, x-, z-, y-solves, divergence + x-transform, z-transform, y-Thomas, inverse z- and
// x-transforms, correction (with the CFL on the last stage), and first the step scalars and
// the stamp-guarded tables.  Channel i of each launch is item i.
template <typename W> struct EnsLaunches {
    MLaunch<KSetStep> set;
    MLaunch<KCnTables<W>> tab;
    MLaunch<KCnStamp> stamp;
    MLaunch<KConv<W, std::is_same<W, float>::value>> conv[3];
    MLaunch<KReduce> red[3];
    MLaunch<KCnX<W>> cnx[3];
    MLaunch<KCnZ<W>> cnz[3];
    MLaunch<KCnY3<W>> cny[3];
    MLaunch<KDivg<W>> divg[3];
    MLaunch<KFftz<W>> fz[3], fzi[3];
    MLaunch<KYtdma<W>> yt[3];
    MLaunch<KFftxInv<W>> fxi[3];
    MLaunch<KUpdate<false>> upd[2];
    MLaunch<KUpdate<true>> updc;
    MLaunch<KReduceMax> rmax;
};

template <typename W>
int ens_make_launches(chs_ens *e, const double *dt, const double *cvisc, EnsLaunches<W> &el)
{
    const int n = (int)e->ch.size();
    const bool mixed = std::is_same<W, float>::value;
    int rc = 0;
    {
        std::vector<KSetStep::Args> a;
        std::vector<typename KCnTables<W>::Args> b;
        std::vector<KCnStamp::Args> d;
        std::vector<dim3> g1, g2;
        for (int i = 0; i < n; ++i) {
            chs_ctx *c = e->ch[i];
            a.push_back({c->dstep, dt[i], cvisc[i]});
            b.push_back({c->geo, cn_dev<W>(c), c->dstep, c->cn_stamp, 1});
            d.push_back({c->dstep, c->cn_stamp, 1, 3});
            g1.push_back(dim3(1));
            g2.push_back(dim3(5, 3));
        }
        if (!rc) rc = mlaunch_make(e, a, g1, dim3(64), 0, el.set);
        if (!rc) rc = mlaunch_make(e, b, g2, dim3(256), 0, el.tab);
        if (!rc) rc = mlaunch_make(e, d, g1, dim3(64), 0, el.stamp);
    }
    for (int ns = 1; ns <= 3 && !rc; ++ns) {
        std::vector<typename KConv<W, std::is_same<W, float>::value>::Args> aconv;
        std::vector<KReduce::Args> ared;
        std::vector<typename KCnX<W>::Args> acnx;
        std::vector<typename KCnZ<W>::Args> acnz;
        std::vector<typename KCnY3<W>::Args> acny;
        std::vector<typename KDivg<W>::Args> adiv;
        std::vector<typename KFftz<W>::Args> afz, afzi;
        std::vector<typename KYtdma<W>::Args> ayt;
        std::vector<typename KFftxInv<W>::Args> afxi;
        std::vector<KUpdate<false>::Args> aupd;
        std::vector<KReduceMax::Args> armax;
        std::vector<dim3> gconv, gred, gcnx, gcnz, gcny, gdiv, gfz, gyt, gupd;
        size_t shx = 0, shz = 0, shcx = 0;
        int Pmax = 1;
        for (int i = 0; i < n; ++i) {
            chs_ctx *c = e->ch[i];
            const CnTab &t = c->cn[ns];
            const Geo &g = c->geo;
            const Sub s = {c->tgam[ns], c->troh[ns], c->talp[ns], c->dstep};
            const int L = c->L, M = c->M, N = c->N, H = c->H;
            W *h0 = (W *)c->h0, *rhs = (W *)c->rhs, *rhs2 = (W *)c->rhs2;
            cpx<W> *spec = (cpx<W> *)c->spec;
            const int KC = c->tune.conv_kc, KZ = c->tune.cnz_kz, KU = c->tune.upd_kz;
            const int KY3 = c->tune.cny3_ky > 0 ? c->tune.cny3_ky : M;
            const int NLX = c->tune.cnx_nlx > 0 ? c->tune.cnx_nlx : ((2048 / L) < 1 ? 1 : 2048 / L);
            const int TK = c->tune.fftx_tk > 0 ? c->tune.fftx_tk : tk_for(L, N);
            const int NL = c->tune.fftz_nl > 0 ? c->tune.fftz_nl : nlz_for(N);
            const PoisArgs pa{c->px, c->pz, c->twx, c->twz};
            const dim3 gc((L + 63) / 64, (M + 3) / 4, (N + KC - 1) / KC);
            aconv.push_back({g, s, c->q, c->pr, h0, rhs, c->partial, KC});
            gconv.push_back(gc);
            ared.push_back({c->partial, (int)(gc.x * gc.y * gc.z), 1.0 / c->volm, c->scal});
            gred.push_back(dim3(1));
            acnx.push_back({g, (const W *)t.gx, t.Dx, t.fullx, rhs, c->scal, NLX});
            gcnx.push_back(dim3((M * N + NLX - 1) / NLX, 3));
            shcx = std::max(shcx, sizeof(W) * NLX * L);
            acnz.push_back({g, (const W *)t.gz, rhs, rhs2, KZ, t.Dz, nullptr});
            gcnz.push_back(dim3((L + 63) / 64, (M + 3) / 4, 3 * ((N + KZ - 1) / KZ)));
            acny.push_back({g, (const W *)t.by[0], (const W *)t.by[1], (const W *)t.by[2], rhs2, c->q, c->qs, KY3,
                            t.Dy[0]});
            gcny.push_back(dim3((L + 63) / 64, (N + 3) / 4, 3 * ((M + KY3 - 1) / KY3)));
            adiv.push_back({g, pa, c->dstep, c->talp[ns], c->qs, spec, TK});
            gdiv.push_back(dim3(M, N / TK));
            shx = std::max(shx, sizeof(cpx<W>) * (L + (TK / 2) * (L + 1)));
            afz.push_back({g, pa, spec, NL, -1});
            afzi.push_back({g, pa, spec, NL, +1});
            gfz.push_back(dim3((M * H + NL - 1) / NL));
            shz = std::max(shz, sizeof(cpx<W>) * (N + NL * (N + 1)));
            const int C = ytdma_chunk_rows(c->tune.ytdma_c, M, H * N), P = (M + C - 1) / C;
            Pmax = std::max(Pmax, P);
            ayt.push_back({g, c->d_amph, c->d_apph, c->pp, mixed ? c->pp32 : (const float *)nullptr, spec, P, C});
            gyt.push_back(dim3((H * N + 63) / 64));
            afxi.push_back({g, pa, (const cpx<W> *)spec, c->phi, TK});
            const dim3 gu((L + 63) / 64, (M + 3) / 4, (N + KU - 1) / KU);
            aupd.push_back({g, s, c->phi, c->qs, c->q, c->pr, KU, c->partial});
            gupd.push_back(gu);
            armax.push_back({c->partial, (int)(gu.x * gu.y * gu.z), c->dstep + 2});
        }
        const int k = ns - 1;
        if (!rc) rc = mlaunch_make(e, aconv, gconv, dim3(64, 4), 0, el.conv[k]);
        if (!rc) rc = mlaunch_make(e, ared, gred, dim3(256), 0, el.red[k]);
        if (!rc) rc = mlaunch_make(e, acnx, gcnx, dim3(256), shcx, el.cnx[k]);
        if (!rc) rc = mlaunch_make(e, acnz, gcnz, dim3(64, 4), 0, el.cnz[k]);
        if (!rc) rc = mlaunch_make(e, acny, gcny, dim3(64, 4), 0, el.cny[k]);
        if (!rc) rc = mlaunch_make(e, adiv, gdiv, dim3(256), shx, el.divg[k]);
        if (!rc) rc = mlaunch_make(e, afz, gfz, dim3(256), shz, el.fz[k]);
        if (!rc) rc = mlaunch_make(e, ayt, gyt, dim3(64, Pmax), 0, el.yt[k]);
        if (!rc) rc = mlaunch_make(e, afzi, gfz, dim3(256), shz, el.fzi[k]);
        if (!rc) rc = mlaunch_make(e, afxi, gdiv, dim3(256), shx, el.fxi[k]);
        if (ns < 3) {
            if (!rc) rc = mlaunch_make(e, aupd, gupd, dim3(64, 4), 0, el.upd[k]);
        } else {
            if (!rc) rc = mlaunch_make(e, aupd, gupd, dim3(64, 4), 0, el.updc);
            if (!rc) rc = mlaunch_make(e, armax, gred, dim3(1024), 0, el.rmax);
        }
    }
    (void)mixed;
    return rc;
}

template <typename W> void ens_record(const EnsLaunches<W> &el, hipStream_t st)
{
    el.set.launch(st);
    el.tab.launch(st);
    el.stamp.launch(st);
    for (int k = 0; k < 3; ++k) {
        el.conv[k].launch(st);
        el.red[k].launch(st);
        el.cnx[k].launch(st);
        el.cnz[k].launch(st);
        el.cny[k].launch(st);
        el.divg[k].launch(st);
        el.fz[k].launch(st);
        el.yt[k].launch(st);
        el.fzi[k].launch(st);
        el.fxi[k].launch(st);
        if (k < 2) el.upd[k].launch(st);
    }
    el.updc.launch(st);
    el.rmax.launch(st);
}

// Capture of the batched step graph into e->gexec (the item tables are made first: their
// uploads cannot be captured).
template <typename W> int ens_capture_batched(chs_ens *e, const double *dt, const double *cvisc)
{
    for (void *b : e->mbufs) (void)hipFree(b);
    e->mbufs.clear();
    EnsLaunches<W> el;
    if (int rc = ens_make_launches<W>(e, dt, cvisc, el)) return rc;
    hipGraph_t graph = nullptr;
    hipError_t he = hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal);
    if (he != hipSuccess) return fail(nullptr, CHS_E_HIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(he));
    ens_record<W>(el, e->stream);
    he = hipStreamEndCapture(e->stream, &graph);
    if (he != hipSuccess) return fail(nullptr, CHS_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(he));
    he = hipGraphInstantiate(&e->gexec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (he != hipSuccess) return fail(nullptr, CHS_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(he));
    return 0;
}

}  // namespace

extern "C" {

int chs_ens_set_batched(chs_ens *e, int batched)
{
    if (!e || (batched != 0 && batched != 1)) return fail(nullptr, CHS_E_ARG, "chs_ens_set_batched: bad argument");
    e->batched = batched == 1;
    return 0;
}

int chs_ens_step(chs_ens *e, const double *dt, const double *cvisc)
{
    if (!e || !dt || !cvisc) return fail(nullptr, CHS_E_ARG, "chs_ens_step: bad argument");
    const int n = (int)e->ch.size();
    // tables first (no-ops when current); a rebuild moves them, so the graph is re-made
    for (int i = 0; i < n; ++i) {
        chs_ctx *c = e->ch[i];
        if (!c->have_fields) return fail(c, CHS_E_STATE, "chs_ens_step: channel without fields");
        if (!(dt[i] > 0) || !(cvisc[i] > 0)) return fail(c, CHS_E_ARG, "dt and cvisc must be positive");
        for (int ns = 1; ns <= 3; ++ns)
            if (int rc = build_cn(c, ns, dt[i], cvisc[i])) return rc;
    }
    const bool batched = e->batched && ens_batchable(e);
    bool same = e->gexec && (int)e->gdt.size() == n && e->gbatched == batched;
    for (int i = 0; same && i < n; ++i)
        same = e->gdt[i] == dt[i] && e->gcv[i] == cvisc[i] && e->gtune[i] == e->ch[i]->tune_gen &&
               e->gcn[i] == e->ch[i]->cn_gen;
    if (!same && batched) {
        if (e->gexec) {
            (void)hipGraphExecDestroy(e->gexec);
            e->gexec = nullptr;
        }
        int rc = e->ch[0]->mode == CHS_MIXED ? ens_capture_batched<float>(e, dt, cvisc)
                                             : ens_capture_batched<double>(e, dt, cvisc);
        if (rc) return rc;
        e->gbatched = true;
        e->gdt.assign(dt, dt + n);
        e->gcv.assign(cvisc, cvisc + n);
        e->gtune.resize(n);
        e->gcn.resize(n);
        for (int i = 0; i < n; ++i) {
            e->gtune[i] = e->ch[i]->tune_gen;
            e->gcn[i] = e->ch[i]->cn_gen;
        }
    } else if (!same) {
        if (e->gexec) {
            (void)hipGraphExecDestroy(e->gexec);
            e->gexec = nullptr;
        }
        e->gbatched = false;
        hipGraph_t graph = nullptr;
        hipError_t he = hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal);
        if (he != hipSuccess) return fail(nullptr, CHS_E_HIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(he));
        int rc = 0;
        // fork: every lane joins the capture behind the origin stream
        const int nl = (int)e->lanes.size();
        hipError_t ce = hipEventRecord(e->ev[0], e->stream);
        for (int s = 0; s < nl && ce == hipSuccess; ++s) ce = hipStreamWaitEvent(e->lanes[s], e->ev[0], 0);
        if (ce != hipSuccess) rc = fail(nullptr, CHS_E_HIP, std::string("chs_ens_step fork: ") + hipGetErrorString(ce));
        // RK stages outermost; channel i on lane i % nl (its stages stay in order on that lane).
        // Each channel's DT, CVISC and (stamp-guarded) tables are set inside the graph.
        for (int ns = 1; ns <= 3 && !rc; ++ns)
            for (int i = 0; i < n && !rc; ++i) {
                chs_ctx *c = e->ch[i];
                c->stream = e->lanes[i % nl];
                if (ns == 1) {
                    c->dv_host = false;
                    set_step(c, dt[i], cvisc[i]);
                    cn_launch(c, 1, 3);
                }
                rc = c->mode == CHS_MIXED ? substep_impl<float>(c, ns, dt[i], cvisc[i], false, ns == 3)
                                          : substep_impl<double>(c, ns, dt[i], cvisc[i], false, ns == 3);
            }
        for (auto *c : e->ch) c->stream = e->stream;
        // join: the origin stream waits for every lane before the capture ends
        for (int s = 0; s < nl && ce == hipSuccess; ++s) {
            ce = hipEventRecord(e->ev[1 + s], e->lanes[s]);
            if (ce == hipSuccess) ce = hipStreamWaitEvent(e->stream, e->ev[1 + s], 0);
        }
        if (ce != hipSuccess && !rc) rc = fail(nullptr, CHS_E_HIP, std::string("chs_ens_step join: ") + hipGetErrorString(ce));
        he = hipStreamEndCapture(e->stream, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (he != hipSuccess) return fail(nullptr, CHS_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(he));
        he = hipGraphInstantiate(&e->gexec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (he != hipSuccess) return fail(nullptr, CHS_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(he));
        e->gdt.assign(dt, dt + n);
        e->gcv.assign(cvisc, cvisc + n);
        e->gtune.resize(n);
        e->gcn.resize(n);
        for (int i = 0; i < n; ++i) {
            e->gtune[i] = e->ch[i]->tune_gen;
            e->gcn[i] = e->ch[i]->cn_gen;
        }
    }
    hipError_t he = hipGraphLaunch(e->gexec, e->stream);
    if (he != hipSuccess) return fail(nullptr, CHS_E_HIP, std::string("hipGraphLaunch: ") + hipGetErrorString(he));
    for (int i = 0; i < n; ++i) {
        chs_ctx *c = e->ch[i];
        c->dv_host = true;
        c->dv_dt = dt[i];
        c->dv_cv = cvisc[i];
        c->cfl_valid = true;
    }
    return 0;
}

int chs_ens_sync(chs_ens *e)
{
    if (!e) return fail(nullptr, CHS_E_ARG, "null ensemble");
    hipError_t he = hipStreamSynchronize(e->stream);
    if (he != hipSuccess) return fail(nullptr, CHS_E_HIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(he));
    return 0;
}

}  // extern "C"

// Stream of a context, for the library's other components that enqueue work behind its
// solves (the main-domain momentum substep uses a context's Poisson solver).
__attribute__((visibility("hidden"))) hipStream_t chs_internal_stream(chs_ctx *c) { return c ? c->stream : nullptr; }

// The device Poisson solve of a plain FP64 context left in the context's own solution array
// (halo layout [N][M+2][L], ghost rows of BC_WALL_DPH), which chs_internal_phi returns: for a
// caller that reads the solution in place instead of a copy.  gauge 1: FISHPACK's (0,0)-mode gauge.
__attribute__((visibility("hidden"))) int chs_internal_poisson_in_place(chs_ctx *c, const double *rhs, int gauge)
{
    if (!c || !rhs || c->mode != CHS_PLAIN_FP64) return fail(c, CHS_E_ARG, "in-place Poisson: bad argument");
    poisson_launch<double>(c, nullptr, 1.0, rhs, gauge);
    HIPCHK(c, hipGetLastError());
    return 0;
}
__attribute__((visibility("hidden"))) const double *chs_internal_phi(chs_ctx *c) { return c ? c->phi : nullptr; }

extern "C" int chs_mean_zx_tg(chs_ctx *c, double *stats, double *scal)
{
    if (!c || (!stats && !scal)) return fail(c, CHS_E_ARG, "chs_mean_zx_tg: null argument");
    const int L = c->L, M = c->M, N = c->N;
    // coefficients (PP_MEAN_ZX.f90:34-58, :261-292): VL1313_tg = 1 / (NCL1 NCL3)
    const double vl = 1.0 / (double)(L * N);
    const double cog1[3][3] = {{vl * c->dxi, vl * 0.25, vl * 0.25 * c->dzi},
                               {vl * 0.25 * c->dxi, vl, vl * 0.25 * c->dzi},
                               {vl * 0.25 * c->dxi, vl * 0.25, vl * c->dzi}};
    std::vector<double> coef(122);
    for (int a = 0; a < 3; ++a) coef[a] = vl * 0.5;
    coef[3] = vl;
    for (int a = 0; a < 3; ++a) coef[4 + a] = vl * 0.5;
    for (int t = 0; t < 6; ++t) coef[7 + t] = vl * (0.5 * 0.5);
    for (int t = 0; t < 10; ++t) coef[13 + t] = vl * (0.5 * 0.5 * 0.5);
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) coef[23 + 3 * a + b] = coef[32 + 3 * a + b] = cog1[a][b];
    for (int l1 = 0; l1 < 9; ++l1)
        for (int l2 = 0; l2 < 9; ++l2) coef[41 + 9 * l1 + l2] = cog1[l1 / 3][l1 % 3] * cog1[l2 / 3][l2 % 3];
    const size_t n = (size_t)M * 122;
    double *buf = nullptr;
    HIPCHK(c, hipMalloc(&buf, sizeof(double) * (n * L + 122 + n)));
    double *part = buf, *dcoef = buf + n * L, *dout = dcoef + 122;
    std::vector<double> hout(n);
    hipError_t e = hipMemcpyAsync(dcoef, coef.data(), sizeof(double) * 122, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        const dim3 grd((L + 63) / 64, M);
        hipLaunchKernelGGL(k_mean_zx<0>, grd, dim3(64), 0, c->stream, c->geo, c->q, c->pr, part);
        hipLaunchKernelGGL(k_mean_zx<1>, grd, dim3(64), 0, c->stream, c->geo, c->q, c->pr, part);
        hipLaunchKernelGGL(k_mean_zx<2>, grd, dim3(64), 0, c->stream, c->geo, c->q, c->pr, part);
        hipLaunchKernelGGL(k_mean_zx<3>, grd, dim3(64), 0, c->stream, c->geo, c->q, c->pr, part);
        hipLaunchKernelGGL(k_mean_zx<4>, grd, dim3(64), 0, c->stream, c->geo, c->q, c->pr, part);
        hipLaunchKernelGGL(k_mean_zx_sum, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, L, (int)n,
                           part, dcoef, dout);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(hout.data(), dout, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(buf);
    if (e != hipSuccess) return fail(c, CHS_E_HIP, std::string("chs_mean_zx_tg: ") + hipGetErrorString(e));
    if (stats) std::copy(hout.begin(), hout.end(), stats);
    if (scal) {   // U1mean_WORK_tg, U1maxx_WORK_tg (PP_MEAN_ZX.f90:294-305), Area_inlet (COORDJR.f90:368-373)
        double um = 0.0, ux = 0.0, area = 0.0;
        for (int J = 1; J <= M; ++J) {
            um = um + hout[(size_t)(J - 1) * 122] / c->dyfi[J - 1];
            ux = std::max(ux, std::fabs(hout[(size_t)(J - 1) * 122]));
        }
        for (int J = 1; J <= M; ++J)
            for (int k = 0; k < N; ++k) area = area + 1.0 / c->dyfi[J - 1] / c->dzi;
        scal[0] = um / c->dzi * (double)N / area;
        scal[1] = ux;
    }
    return 0;
}
