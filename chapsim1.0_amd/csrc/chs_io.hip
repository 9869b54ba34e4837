// Main-domain momentum substep with variable density and viscosity (SOLVERRK3_MOM_io,
// SOLVERRK3_MOM_IO.f90:11-120, explicit or Crank-Nicolson viscous scheme) and the staggered
// properties SOLVE
// computes before it in a heated run (DENSITY_TIME_DERIVATION, DENSITY_Staggered, MU_Staggered;
// SOLVE.f90:122-124), for the periodic plane channel (RCCI1 = RCCI2 = RNDI1 = 1, no inlet,
// iWeightedPre = 0).  The pressure Poisson solve is the FFT99 one of a chs_ctx on the same
// mesh (FFT99_POIS3D_periodicxz(IIO), the reference's is_FFT_FFT99 option).
//
// Every expression keeps the reference's operation order with FMA contraction off, so the
// results follow the reference's rounding.  State stays in HBM; fields use the reference's
// layout: halo fields [N][M+2][L] (x fastest, y rows 0..M+1), cell fields [N][M][L].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <string>
#include <utility>
#include <vector>

#include "../../include/chapsim_hip.h"

hipStream_t chs_internal_stream(chs_ctx *c);
int chs_internal_poisson_in_place(chs_ctx *c, const double *rhs, int gauge);
const double *chs_internal_phi(chs_ctx *c);
void chs_internal_eng_fields(chs_eng *e, hipStream_t *s, double **d0, double **m0, double **m, double **g);
void chs_internal_eng_set_g(chs_eng *e, const double *g);
double *chs_internal_eng_density(chs_eng *e);

namespace {

// Maximum that propagates NaN (fmax drops a NaN operand, so a diverged field would report a
// finite CFL or divergence).
__device__ __forceinline__ double nanmax(double a, double b) { return (a > b || a != a) ? a : b; }

thread_local std::string g_io_err;

struct IoDev {
    int L, M, N;
    double dxi, dzi, cvisc, f_a;
    int gravity;
    const double *dyfi, *dyci, *wff, *wfb;    // indexed by the reference's J (dyfi 1..M, others 1..M+1)
    double *g, *g0, *q, *dstg, *mstg;         // 3 halo fields each
    double *pr, *d0, *m0;                     // halo fields
    double *explt0, *h;                       // 3 cell fields each
    float *explt0f, *hf;                      // the same in the mixed form (work precision FP32)
    double *drho, *divu, *rhsp;               // cell fields
    const double *dph;                        // DPH_io: the Poisson context's solution, halo field
    double *partial, *scal;
    // Crank-Nicolson scheme (iVisScheme = 1)
    double *m1;                               // Viscousity, the current viscosity (halo field)
    double *work;                             // line-solve recurrences: 3 x 3 cell fields
    const double *amvr, *apvr;                // [3][M+2] AMVR / APVR(J, IDR) at [IDR-1][J]
    double amvr1, apvrn, dxqi, dzqi;
};

__device__ __forceinline__ size_t hx(const IoDev &d, int i, int J, int k) { return ((size_t)k * (d.M + 2) + J) * d.L + i; }
__device__ __forceinline__ size_t cx(const IoDev &d, int i, int J, int k) { return ((size_t)k * d.M + (J - 1)) * d.L + i; }
__device__ __forceinline__ size_t hsz(const IoDev &d) { return (size_t)d.N * (d.M + 2) * d.L; }
__device__ __forceinline__ size_t csz(const IoDev &d) { return (size_t)d.N * d.M * d.L; }

// EXPLT0 and the increment in the work precision W: double (plain) or float (mixed form).
template <typename W> __device__ __forceinline__ W *ioh(const IoDev &d);
template <> __device__ __forceinline__ double *ioh<double>(const IoDev &d) { return d.h; }
template <> __device__ __forceinline__ float *ioh<float>(const IoDev &d) { return d.hf; }
template <typename W> __device__ __forceinline__ W *ioex(const IoDev &d);
template <> __device__ __forceinline__ double *ioex<double>(const IoDev &d) { return d.explt0; }
template <> __device__ __forceinline__ float *ioex<float>(const IoDev &d) { return d.explt0f; }

// Tile width of the staged planes: columns i0-2 .. i0+65 (the halo column each side and one
// more on the left, so that a tile row is 34 aligned pairs of columns, one 16-byte load each).
constexpr int IO_TW = 68;

// Tile of launch slot p of nb: the hardware deals slots to the XCDs in turn (slot p to XCD
// p mod nxcd); each XCD gets a contiguous run of tiles, so that the halo rows a tile shares
// with the tiles above and below it are read through the same L2.
__device__ __forceinline__ int io_xcd_tile(int p, int nb, int nxcd)
{
    if (nxcd <= 1) return p;
    const int xcd = p % nxcd, slot = p / nxcd, per = nb / nxcd, rem = nb % nxcd;
    return xcd < rem ? xcd * (per + 1) + slot : rem * (per + 1) + (xcd - rem) * per + slot;
}


// Thread (i, J = j0 + row, k) of a 64 x 4 block on a (ceil(L/64), ceil(rows/4), N) grid.
#define IO_THREAD(j0)                                        \
    const int i = blockIdx.x * 64 + threadIdx.x;             \
    const int J = (j0) + (int)(blockIdx.y * 4 + threadIdx.y); \
    const int k = blockIdx.z;                                \
    const int L = d.L, M = d.M, N = d.N;                     \
    (void)N;

// DENSITY_TIME_DERIVATION (DIVG.f90:249-317), DENSITY_Staggered (MASSFLUX_CALC_IO.f90:262-323)
// and MU_Staggered (:175-259), rows J = 1..M+1, and the velocity the next momentum substep
// starts from (VELOCITY_CALC_io with BC_WALL_Q_io, as k_io_velocity), which needs only this
// point's G and staggered density.  Each thread marches KC z-planes: the values plane k+1
// needs from plane k (D and V at (i, J), (i-1, J), (i, J-1); G_z and G0_z at (i, J)) stay in
// registers, so a plane costs 8 state loads instead of 14.
// Plane-k fields of the staggered-property point: X(f, di, dj) = f at (i + di, J + dj, k).
enum IoPF { PD, PV, PGX, PG0X, PGY, PG0Y };

// One point (i, J, k), J = 1..M+1: the plane-(k-1) values (D, V at (i, J); V at (i-1, J) and
// (i, J-1)) and G_z, G0_z at (i, J) of planes k and k+1 come as values.
// Row J's mesh coefficients of the staggered-property point: DYFI(min(J, M)), YCL2ND_WFF/WFB(J).
struct IoPropRow {
    double coe2, wff, wfb;
};
__device__ __forceinline__ IoPropRow io_prop_row(const IoDev &d, int J)
{
    return {d.dyfi[min(J, d.M)], d.wff[J], d.wfb[J]};
}

template <typename A>
__device__ __forceinline__ void io_props_point(const IoDev &d, const A &X, const IoPropRow &rc, int i, int J, int k,
                                               double tg, double tr, double Dm, double Vm, double Vmim, double Vmb,
                                               double g2c, double g2p, double g02c, double g02p)
{
#pragma clang fp contract(off)
    const int M = d.M;
    const size_t hs = hsz(d), c0 = hx(d, i, J, k);
    const double h = 0.5;
    const double coe2 = rc.coe2, coe3 = d.dzi, wff = rc.wff, wfb = rc.wfb;
    const double Dc = X(PD, 0, 0), Vc = X(PV, 0, 0), Vim = X(PV, -1, 0);
    const double Vb = X(PV, 0, -1), Vimb = X(PV, -1, -1);
    double gx = 0.0, gy = 0.0, gz = g2c, dsx = 0.0, dsy, dsz = 0.0;
    if (J <= M) {
        gx = X(PGX, 0, 0);
        gy = X(PGY, 0, 0);
        const double divx = (X(PGX, 1, 0) - gx) * d.dxi;
        const double divx0 = (X(PG0X, 1, 0) - X(PG0X, 0, 0)) * d.dxi;
        const double divy = (X(PGY, 0, 1) - gy) * coe2;
        const double divy0 = (X(PG0Y, 0, 1) - X(PG0Y, 0, 0)) * coe2;
        const double divz = (g2p - g2c) * coe3;
        const double divz0 = (g02p - g02c) * coe3;
        d.drho[cx(d, i, J, k)] = -tg * (divx + divy + divz) - tr * (divx0 + divy0 + divz0);
        dsx = (Dc + X(PD, -1, 0)) * h;
        dsz = (Dc + Dm) * h;
        d.dstg[c0] = dsx;
        d.dstg[2 * hs + c0] = dsz;
        d.mstg[hs + c0] = (Vc + Vim + Vm + Vmim) * h * h;
    }
    if (J == 1) {
        dsy = X(PD, 0, -1);
        d.dstg[hs + c0] = dsy;
        d.mstg[c0] = Vb;
        d.mstg[2 * hs + c0] = Vb;
    } else if (J == M + 1) {
        dsy = Dc;
        d.dstg[hs + c0] = dsy;
        d.mstg[c0] = Vc;
        d.mstg[2 * hs + c0] = Vc;
    } else {
        dsy = wff * Dc + wfb * X(PD, 0, -1);
        d.dstg[hs + c0] = dsy;
        d.mstg[c0] = ((Vc + Vim) * wff + (Vb + Vimb) * wfb) * h;
        d.mstg[2 * hs + c0] = ((Vm + Vc) * wff + (Vmb + Vb) * wfb) * h;
    }
    // the velocity the next momentum substep starts from (VELOCITY_CALC_io, BC_WALL_Q_io)
    d.q[c0] = J == M + 1 ? 0.0 : gx / dsx;
    d.q[hs + c0] = (J == M + 1 || J == 1) ? 0.0 : gy / dsy;
    d.q[2 * hs + c0] = J == M + 1 ? 0.0 : gz / dsz;
    if (J == 1)
        for (int c = 0; c < 3; ++c) d.q[c * hs + hx(d, i, 0, k)] = 0.0;
}

__global__ __launch_bounds__(256) void k_io_props_z(IoDev d, double tg, double tr, int KC)
{
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int J = (int)(blockIdx.y * 4 + threadIdx.y) + 1;
    const int L = d.L, M = d.M, N = d.N;
    if (i >= L || J > M + 1) return;
    const int k0 = blockIdx.z * KC, k1 = min(N, k0 + KC);
    const int ip = i == L - 1 ? 0 : i + 1, im = i == 0 ? L - 1 : i - 1;
    const size_t hs = hsz(d);
    const double *D = d.d0, *V = d.m0, *G = d.g, *G0 = d.g0;
    const IoPropRow rc = io_prop_row(d, J);
    const int kb = k0 == 0 ? N - 1 : k0 - 1;
    double Dm = D[hx(d, i, J, kb)], Vm = V[hx(d, i, J, kb)], Vmim = V[hx(d, im, J, kb)];
    double Vmb = V[hx(d, i, J - 1, kb)];
    double g2c = 0.0, g02c = 0.0;
    if (J <= M) {
        g2c = G[2 * hs + hx(d, i, J, k0)];
        g02c = G0[2 * hs + hx(d, i, J, k0)];
    }
    for (int k = k0; k < k1; ++k) {
        const int kp = k == N - 1 ? 0 : k + 1;
        double g2p = 0.0, g02p = 0.0;
        if (J <= M) {
            g2p = G[2 * hs + hx(d, i, J, kp)];
            g02p = G0[2 * hs + hx(d, i, J, kp)];
        }
        auto X = [&](int f, int di, int dj) {
            const double *a = f == PD ? D : f == PV ? V : f == PGX ? G : f == PG0X ? G0 : f == PGY ? G + hs : G0 + hs;
            return a[hx(d, di > 0 ? ip : (di < 0 ? im : i), J + dj, k)];
        };
        io_props_point(d, X, rc, i, J, k, tg, tr, Dm, Vm, Vmim, Vmb, g2c, g2p, g02c, g02p);
        Dm = X(PD, 0, 0);
        Vm = X(PV, 0, 0);
        Vmim = X(PV, -1, 0);
        Vmb = X(PV, 0, -1);
        g2c = g2p;
        g02c = g02p;
    }
}

// The same from LDS: a 64 x 4 block owns columns i0..i0+63 and rows J0..J0+3 (J <= M+1) and
// marches KC z-planes; D, V and the x and y components of G and G0 are staged per plane as
// 6 x 68 tiles (columns i0-2 .. i0+65, loaded as 16-byte column pairs), plane k+1's tiles
// loaded while plane k is computed; the plane-(k-1) values stay in registers, G_z and G0_z of
// the point are loaded a plane ahead.  Identical arithmetic to k_io_props_z.
__global__ __launch_bounds__(256) void k_io_props_lds(IoDev d, double tg, double tr, int KC, int nxcd)
{
    constexpr int BY = 4, NT = 64 * BY, TW = IO_TW, TS = (BY + 2) * TW, NF = 6, NCH = TS / 2;
    static_assert(NCH <= NT, "one column pair per thread");
    __shared__ __align__(16) double tiles[NF * TS];
    const int L = d.L, M = d.M, N = d.N;
    const int tx = threadIdx.x, ty = __builtin_amdgcn_readfirstlane(threadIdx.y), t = ty * 64 + tx;
    const int gx = gridDim.x, gy = gridDim.y;
    const int tile = io_xcd_tile(blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z), gx * gy * gridDim.z, nxcd);
    const int bx = tile % gx, by = (tile / gx) % gy, bz = tile / (gx * gy);
    const int i0 = bx * 64, J0 = 1 + by * BY, k0 = bz * KC, k1 = min(N, k0 + KC);
    const int i = i0 + tx, J = J0 + ty;
    const bool act = i < L && J <= M + 1;
    const size_t hs = hsz(d), ph = (size_t)(M + 2) * L;
    const bool stg = t < NCH;
    unsigned oh;
    {
        const int n = stg ? t : 0, r = n / (TW / 2), q = n % (TW / 2), jj = J0 - 1 + r;
        oh = (unsigned)(min(jj, M + 1) * L + ((i0 - 2 + 2 * q) % L + L) % L);
    }
    const double *base[NF] = {d.d0, d.m0, d.g, d.g0, d.g + hs, d.g0 + hs};
    double pre0[NF], pre1[NF];
    auto load = [&](int kk) {
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            const double2 v = *reinterpret_cast<const double2 *>(base[f] + (size_t)kk * ph + oh);
            pre0[f] = v.x;
            pre1[f] = v.y;
        }
    };
    auto store = [&]() {
        if (stg) {
#pragma unroll
            for (int f = 0; f < NF; ++f)
                reinterpret_cast<double2 *>(&tiles[f * TS])[t] = make_double2(pre0[f], pre1[f]);
        }
    };
    const int ir = act ? i : 0, Jr = act ? J : 1, imr = ir == 0 ? L - 1 : ir - 1;
    const bool gz = act && J <= M;   // rows with a z mass flux
    const IoPropRow rc = io_prop_row(d, act ? J : 1);
    const int kb = k0 == 0 ? N - 1 : k0 - 1;
    double Dm = d.d0[hx(d, ir, Jr, kb)], Vm = d.m0[hx(d, ir, Jr, kb)], Vmim = d.m0[hx(d, imr, Jr, kb)];
    double Vmb = d.m0[hx(d, ir, Jr - 1, kb)];
    double g2c = 0.0, g02c = 0.0, g2p = 0.0, g02p = 0.0, g2n = 0.0, g02n = 0.0;
    if (gz) {
        const int kp = k0 == N - 1 ? 0 : k0 + 1;
        g2c = d.g[2 * hs + hx(d, i, J, k0)];
        g02c = d.g0[2 * hs + hx(d, i, J, k0)];
        g2p = d.g[2 * hs + hx(d, i, J, kp)];
        g02p = d.g0[2 * hs + hx(d, i, J, kp)];
    }
    load(k0);
    store();
    __syncthreads();
    const int cell = (ty + 1) * TW + tx + 2;
    for (int k = k0; k < k1; ++k) {
        const bool more = k + 1 < k1;
        if (more) {
            if (gz) {   // G_z, G0_z of plane k+2: the next plane's k+1 values
                const int kn = k + 2 >= N ? k + 2 - N : k + 2;
                g2n = d.g[2 * hs + hx(d, i, J, kn)];
                g02n = d.g0[2 * hs + hx(d, i, J, kn)];
            }
            load(k + 1 == N ? 0 : k + 1);
        }
        auto X = [&](int f, int di, int dj) { return tiles[f * TS + cell + dj * TW + di]; };
        if (act) io_props_point(d, X, rc, i, J, k, tg, tr, Dm, Vm, Vmim, Vmb, g2c, g2p, g02c, g02p);
        Dm = X(PD, 0, 0);
        Vm = X(PV, 0, 0);
        Vmim = X(PV, -1, 0);
        Vmb = X(PV, 0, -1);
        g2c = g2p;
        g02c = g02p;
        g2p = g2n;
        g02p = g02n;
        __syncthreads();
        if (more) store();
        __syncthreads();
    }
}

// VELOCITY_CALC_io (MASSFLUX_CALC_IO.f90:55-107) with BC_WALL_Q_io: rows J = 0..M+1.
__global__ __launch_bounds__(256) void k_io_velocity(IoDev d, const double *__restrict__ gs)
{
#pragma clang fp contract(off)
    IO_THREAD(0)
    if (i >= L || J > M + 1) return;
    const size_t hs = hsz(d), c0 = hx(d, i, J, k);
    for (int c = 0; c < 3; ++c) {
        const size_t o = c * hs + c0;
        d.q[o] = (J == 0 || J == M + 1 || (c == 1 && J == 1)) ? 0.0 : gs[o] / d.dstg[o];
    }
}

// DIVG_U_io (DIVG.f90:140-246): rows 1..M.
__global__ __launch_bounds__(256) void k_io_divu(IoDev d)
{
#pragma clang fp contract(off)
    IO_THREAD(1)
    if (i >= L || J > M) return;
    const int ip = i == L - 1 ? 0 : i + 1, kp = k == N - 1 ? 0 : k + 1;
    const size_t hs = hsz(d), c0 = hx(d, i, J, k);
    const double *Q = d.q;
    const double dx = (Q[hx(d, ip, J, k)] - Q[c0]) * d.dxi;
    const double dy = (Q[hs + hx(d, i, J + 1, k)] - Q[hs + c0]) * d.dyfi[J];
    const double dz = (Q[2 * hs + hx(d, i, J, kp)] - Q[2 * hs + c0]) * d.dzi;
    d.divu[cx(d, i, J, k)] = (dx + dy + dz) / 3.0;
}

// Block sum (256 threads) in a fixed order; thread 0 gets the result.
__device__ double block_sum256(double v, double *sh)
{
    const int t = threadIdx.x + 64 * threadIdx.y;
    sh[t] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) sh[t] = sh[t] + sh[t + s];
        __syncthreads();
    }
    return sh[0];
}

// Fields the momentum stencil reads: G (the previous mass flux, G0_io), Q, the staggered
// viscosities MU_STG (x-y, x-z, y-z), the viscosity (Viscousity0 for the explicit scheme,
// the current Viscousity for Crank-Nicolson), the pressure, and the explicit scheme's
// divergence (a cell field, rows 1..M).
enum IoField { FG0, FG1, FG2, FQ0, FQ1, FQ2, FMS0, FMS1, FMS2, FM0, FM1, FPR, FDU };

// Mesh coefficients of row J the stencil reads: DYFI(J), DYFI(J-1), DYCI(J), DYCI(J+1),
// YCL2ND_WFF/WFB(J) and (J+1).
enum IoRowC { RDYFI, RDYFIM, RDYCI, RDYCIP, RWFF, RWFB, RWFFP, RWFBP, IO_NROW };

// ac(f, di, dj, dk) = field f at (i + di, J + dj, k + dk), x and z periodic, read from HBM;
// EXPLT0 and the increment in the work precision W.
template <typename W = double> struct IoAtHbm {
    const IoDev &d;
    const double *gs;
    int i, J, k, ip, im, kp, km;
    __device__ __forceinline__ double operator()(int f, int di, int dj, int dk) const
    {
        const int ii = di > 0 ? ip : (di < 0 ? im : i), kk = dk > 0 ? kp : (dk < 0 ? km : k), jj = J + dj;
        const size_t hs = hsz(d), o = hx(d, ii, jj, kk);
        switch (f) {
        case FG0: return gs[o];
        case FG1: return gs[hs + o];
        case FG2: return gs[2 * hs + o];
        case FQ0: return d.q[o];
        case FQ1: return d.q[hs + o];
        case FQ2: return d.q[2 * hs + o];
        case FMS0: return d.mstg[o];
        case FMS1: return d.mstg[hs + o];
        case FMS2: return d.mstg[2 * hs + o];
        case FM0: return d.m0[o];
        case FM1: return d.m1[o];
        case FPR: return d.pr[o];
        default: return d.divu[cx(d, ii, jj, kk)];
        }
    }
    __device__ __forceinline__ double rw(int c) const
    {
        switch (c) {
        case RDYFI: return d.dyfi[J];
        case RDYFIM: return d.dyfi[J - 1];
        case RDYCI: return d.dyci[J];
        case RDYCIP: return d.dyci[J + 1];
        case RWFF: return d.wff[J];
        case RWFB: return d.wfb[J];
        case RWFFP: return d.wff[J + 1];
        default: return d.wfb[J + 1];
        }
    }
    // EXPLT0 of component c and the staggered density of component c at the point
    __device__ __forceinline__ double ex(int c) const { return ioex<W>(d)[(size_t)c * csz(d) + cx(d, i, J, k)]; }
    // the new EXPLT0 and the increment of component c at the point
    __device__ __forceinline__ void put(int c, double acc, double r) const
    {
        const size_t o = (size_t)c * csz(d) + cx(d, i, J, k);
        ioex<W>(d)[o] = (W)acc;
        ioh<W>(d)[o] = (W)r;
    }
    __device__ __forceinline__ double ds(int c) const { return d.dstg[(size_t)c * hsz(d) + hx(d, i, J, k)]; }
};

// One momentum component C (0 x, 1 y, 2 z) at a point: convection (CONVECTION_X/Y/Z_io),
// the viscous terms and RHS_MOM_EXPLICIT_io(NS, C+1): RK combination, EXPLT0 update, pressure
// gradient, buoyancy.  Explicit scheme (CN false): the full variable-mu stress divergence
// (VISCOUS_ALL_EXPLT_X/Y/Z_io, channel branch) joins the convection term.  Crank-Nicolson
// (CN true): the split of VISCOUS_PAR_EXPLT_X/Y/Z_io (VISCOUS_X_EXPLICIT_io.f90:131-293,
// VISCOUS_Y_EXPLICIT_io.f90:278-434, VISCOUS_Z_EXPLICIT_io.f90:300-462) with the current
// viscosity: its explicit part joins the convection term (EX + conv), its implicit part enters
// the increment with alpha dt (RHS_MOM_EXPLICIT_io.f90:87-99).  Their density-correction parts
// multiply DRHOI_STG, which DENSITY_IMPLICIT_add evaluates as 1/D_STG minus the reciprocal of
// D_STG's own expression of the same DENSITY0 (MASSFLUX_CALC_IO.f90:332-416): exactly zero.
// Rows 1..M (2..M for v).  Returns RHS / DYFI for C = 0 (the mass-flux partial), 0 otherwise.
template <int C, bool CN, typename Acc>
__device__ __forceinline__ double io_mom_point(const IoDev &d, const Acc &ac, int i, int J, int k, double tg,
                                               double tr, double ta, double dt)
{
#pragma clang fp contract(off)
    {
        const double hf = 0.5, cv = d.cvisc, dxi = d.dxi, dzi = d.dzi;
        const double wffp = ac.rw(RWFFP), wfbp = ac.rw(RWFBP), wff = ac.rw(RWFF), wfb = ac.rw(RWFB);
        double acc, pgm, imp = 0.0;
        (void)imp;
        if constexpr (C == 0) {
            const double c1 = dxi * hf * hf, c2 = ac.rw(RDYFI) * hf, c3 = dzi * hf * hf;
            double f = (ac(FG0, 1, 0, 0) + ac(FG0, 0, 0, 0)) * (ac(FQ0, 1, 0, 0) + ac(FQ0, 0, 0, 0));
            double b = (ac(FG0, 0, 0, 0) + ac(FG0, -1, 0, 0)) * (ac(FQ0, 0, 0, 0) + ac(FQ0, -1, 0, 0));
            const double h11 = (f - b) * c1;
            f = (ac(FG1, 0, 1, 0) + ac(FG1, -1, 1, 0)) * (wffp * ac(FQ0, 0, 1, 0) + wfbp * ac(FQ0, 0, 0, 0));
            b = (ac(FG1, 0, 0, 0) + ac(FG1, -1, 0, 0)) * (wff * ac(FQ0, 0, 0, 0) + wfb * ac(FQ0, 0, -1, 0));
            const double h12 = (f - b) * c2;
            f = (ac(FG2, 0, 0, 1) + ac(FG2, -1, 0, 1)) * (ac(FQ0, 0, 0, 1) + ac(FQ0, 0, 0, 0));
            b = (ac(FG2, 0, 0, 0) + ac(FG2, -1, 0, 0)) * (ac(FQ0, 0, 0, 0) + ac(FQ0, 0, 0, -1));
            const double h13 = (f - b) * c3;
            acc = -(h11 + h12 + h13);
            if constexpr (CN) {
                const double e1 = -2.0 / 3.0 * dxi * cv, e2 = ac.rw(RDYFI) * cv, e3 = dzi * cv;
                double dudx = (ac(FQ0, 1, 0, 0) - ac(FQ0, 0, 0, 0)) * dxi;
                double dvdy = (ac(FQ1, 0, 1, 0) - ac(FQ1, 0, 0, 0)) * ac.rw(RDYFI);
                double dwdz = (ac(FQ2, 0, 0, 1) - ac(FQ2, 0, 0, 0)) * dzi;
                const double t11fe = ac(FM1, 0, 0, 0) * (dvdy + dwdz), t11fi = ac(FM1, 0, 0, 0) * dudx;
                dudx = (ac(FQ0, 0, 0, 0) - ac(FQ0, -1, 0, 0)) * dxi;
                dvdy = (ac(FQ1, -1, 1, 0) - ac(FQ1, -1, 0, 0)) * ac.rw(RDYFI);
                dwdz = (ac(FQ2, -1, 0, 1) - ac(FQ2, -1, 0, 0)) * dzi;
                const double t11be = ac(FM1, -1, 0, 0) * (dvdy + dwdz), t11bi = ac(FM1, -1, 0, 0) * dudx;
                const double d11e = (t11fe - t11be) * e1, d11i = (t11fi - t11bi) * e1 * (-2.0);
                double vis = ac(FMS0, 0, 1, 0);
                double dvdx = (ac(FQ1, 0, 1, 0) - ac(FQ1, -1, 1, 0)) * dxi;
                double dudy = (ac(FQ0, 0, 1, 0) - ac(FQ0, 0, 0, 0)) * ac.rw(RDYCIP);
                const double t12fe = vis * dvdx, t12fi = vis * dudy;
                vis = ac(FMS0, 0, 0, 0);
                dudy = (ac(FQ0, 0, 0, 0) - ac(FQ0, 0, -1, 0)) * ac.rw(RDYCI);
                dvdx = (ac(FQ1, 0, 0, 0) - ac(FQ1, -1, 0, 0)) * dxi;
                const double t12be = vis * dvdx, t12bi = vis * dudy;
                const double d12e = (t12fe - t12be) * e2, d12i = (t12fi - t12bi) * e2;
                vis = ac(FMS1, 0, 0, 1);
                double dudz = (ac(FQ0, 0, 0, 1) - ac(FQ0, 0, 0, 0)) * dzi;
                double dwdx = (ac(FQ2, 0, 0, 1) - ac(FQ2, -1, 0, 1)) * dxi;
                const double t13fe = vis * dwdx, t13fi = vis * dudz;
                vis = ac(FMS1, 0, 0, 0);
                dudz = (ac(FQ0, 0, 0, 0) - ac(FQ0, 0, 0, -1)) * dzi;
                dwdx = (ac(FQ2, 0, 0, 0) - ac(FQ2, -1, 0, 0)) * dxi;
                const double t13be = vis * dwdx, t13bi = vis * dudz;
                const double d13e = (t13fe - t13be) * e3, d13i = (t13fi - t13bi) * e3;
                acc = d11e + d12e + d13e + acc;
                imp = d11i + d12i + d13i;
            } else {
            const double v1 = 2.0 * dxi * cv, v2 = ac.rw(RDYFI) * cv, v3 = dzi * cv;
            double dudx = (ac(FQ0, 1, 0, 0) - ac(FQ0, 0, 0, 0)) * dxi;
            const double t11f = 0.5 * (ac(FM0, 0, 0, 0) + ac(FM0, 0, 0, 0)) * (dudx - ac(FDU, 0, 0, 0));
            dudx = (ac(FQ0, 0, 0, 0) - ac(FQ0, -1, 0, 0)) * dxi;
            const double t11b = 0.5 * (ac(FM0, -1, 0, 0) + ac(FM0, -1, 0, 0)) * (dudx - ac(FDU, -1, 0, 0));
            const double d11 = (t11f - t11b) * v1;
            double dudy = (ac(FQ0, 0, 1, 0) - ac(FQ0, 0, 0, 0)) * ac.rw(RDYCIP);
            double dvdx = (ac(FQ1, 0, 1, 0) - ac(FQ1, -1, 1, 0)) * dxi;
            const double t12f = ac(FMS0, 0, 1, 0) * (dudy + dvdx);
            dudy = (ac(FQ0, 0, 0, 0) - ac(FQ0, 0, -1, 0)) * ac.rw(RDYCI);
            dvdx = (ac(FQ1, 0, 0, 0) - ac(FQ1, -1, 0, 0)) * dxi;
            const double t12b = ac(FMS0, 0, 0, 0) * (dudy + dvdx);
            const double d12 = (t12f - t12b) * v2;
            double dudz = (ac(FQ0, 0, 0, 1) - ac(FQ0, 0, 0, 0)) * dzi;
            double dwdx = (ac(FQ2, 0, 0, 1) - ac(FQ2, -1, 0, 1)) * dxi;
            const double t13f = ac(FMS1, 0, 0, 1) * (dudz + dwdx);
            dudz = (ac(FQ0, 0, 0, 0) - ac(FQ0, 0, 0, -1)) * dzi;
            dwdx = (ac(FQ2, 0, 0, 0) - ac(FQ2, -1, 0, 0)) * dxi;
            const double t13b = ac(FMS1, 0, 0, 0) * (dudz + dwdx);
            const double d13 = (t13f - t13b) * v3;
            acc = acc + d11 + d12 + d13;
            }
            pgm = (ac(FPR, 0, 0, 0) - ac(FPR, -1, 0, 0)) * (ta * dt * dxi);
        } else if constexpr (C == 1) {
            const double c1 = dxi * hf, c2 = ac.rw(RDYCI) * hf * hf, c3 = dzi * hf;
            double f = (wff * ac(FG0, 1, 0, 0) + wfb * ac(FG0, 1, -1, 0)) * (ac(FQ1, 1, 0, 0) + ac(FQ1, 0, 0, 0));
            double b = (wff * ac(FG0, 0, 0, 0) + wfb * ac(FG0, 0, -1, 0)) * (ac(FQ1, 0, 0, 0) + ac(FQ1, -1, 0, 0));
            const double h21 = (f - b) * c1;
            f = (ac(FG1, 0, 1, 0) + ac(FG1, 0, 0, 0)) * (ac(FQ1, 0, 1, 0) + ac(FQ1, 0, 0, 0));
            b = (ac(FG1, 0, 0, 0) + ac(FG1, 0, -1, 0)) * (ac(FQ1, 0, 0, 0) + ac(FQ1, 0, -1, 0));
            const double h22 = (f - b) * c2;
            f = (wff * ac(FG2, 0, 0, 1) + wfb * ac(FG2, 0, -1, 1)) * (ac(FQ1, 0, 0, 1) + ac(FQ1, 0, 0, 0));
            b = (wff * ac(FG2, 0, 0, 0) + wfb * ac(FG2, 0, -1, 0)) * (ac(FQ1, 0, 0, 0) + ac(FQ1, 0, 0, -1));
            const double h23 = (f - b) * c3;
            acc = -(h21 + h22 + h23);
            if constexpr (CN) {
                const double e1 = dxi * cv, e3 = dzi * cv, e2 = -2.0 / 3.0 * ac.rw(RDYCI) * cv;
                double dudx = (ac(FQ0, 1, 0, 0) - ac(FQ0, 0, 0, 0)) * dxi;
                double dvdy = (ac(FQ1, 0, 1, 0) - ac(FQ1, 0, 0, 0)) * ac.rw(RDYFI);
                double dwdz = (ac(FQ2, 0, 0, 1) - ac(FQ2, 0, 0, 0)) * dzi;
                const double t22fe = ac(FM1, 0, 0, 0) * (dudx + dwdz), t22fi = ac(FM1, 0, 0, 0) * dvdy;
                dudx = (ac(FQ0, 1, -1, 0) - ac(FQ0, 0, -1, 0)) * dxi;
                dvdy = (ac(FQ1, 0, 0, 0) - ac(FQ1, 0, -1, 0)) * ac.rw(RDYFIM);
                dwdz = (ac(FQ2, 0, -1, 1) - ac(FQ2, 0, -1, 0)) * dzi;
                const double t22be = ac(FM1, 0, -1, 0) * (dudx + dwdz), t22bi = ac(FM1, 0, -1, 0) * dvdy;
                const double d22e = (t22fe - t22be) * e2, d22i = (t22fi - t22bi) * e2 * (-2.0);
                double vis = ac(FMS0, 1, 0, 0);
                double dvdx = (ac(FQ1, 1, 0, 0) - ac(FQ1, 0, 0, 0)) * dxi;
                double dudy = (ac(FQ0, 1, 0, 0) - ac(FQ0, 1, -1, 0)) * ac.rw(RDYCI);
                const double t21fe = vis * dudy, t21fi = vis * dvdx;
                vis = ac(FMS0, 0, 0, 0);
                dvdx = (ac(FQ1, 0, 0, 0) - ac(FQ1, -1, 0, 0)) * dxi;
                dudy = (ac(FQ0, 0, 0, 0) - ac(FQ0, 0, -1, 0)) * ac.rw(RDYCI);
                const double t21be = vis * dudy, t21bi = vis * dvdx;
                const double d21e = (t21fe - t21be) * e1, d21i = (t21fi - t21bi) * e1;
                vis = ac(FMS2, 0, 0, 1);
                double dvdz = (ac(FQ1, 0, 0, 1) - ac(FQ1, 0, 0, 0)) * dzi;
                double dwdy = (ac(FQ2, 0, 0, 1) - ac(FQ2, 0, -1, 1)) * ac.rw(RDYCI);
                const double t23fe = vis * dwdy, t23fi = vis * dvdz;
                vis = ac(FMS2, 0, 0, 0);
                dvdz = (ac(FQ1, 0, 0, 0) - ac(FQ1, 0, 0, -1)) * dzi;
                dwdy = (ac(FQ2, 0, 0, 0) - ac(FQ2, 0, -1, 0)) * ac.rw(RDYCI);
                const double t23be = vis * dwdy, t23bi = vis * dvdz;
                const double d23e = (t23fe - t23be) * e3, d23i = (t23fi - t23bi) * e3;
                acc = d21e + d22e + d23e + acc;
                imp = d21i + d22i + d23i;
            } else {
            const double v1 = dxi * cv, v2 = 2.0 * ac.rw(RDYCI) * cv, v3 = dzi * cv;
            double dvdy = (ac(FQ1, 0, 1, 0) - ac(FQ1, 0, 0, 0)) * ac.rw(RDYFI);
            const double t22f = 0.5 * (ac(FM0, 0, 0, 0) + ac(FM0, 0, 0, 0)) * (dvdy - ac(FDU, 0, 0, 0));
            dvdy = (ac(FQ1, 0, 0, 0) - ac(FQ1, 0, -1, 0)) * ac.rw(RDYFIM);
            const double t22b = 0.5 * (ac(FM0, 0, -1, 0) + ac(FM0, 0, -1, 0)) * (dvdy - ac(FDU, 0, -1, 0));
            const double d22 = (t22f - t22b) * v2;
            double dvdx = (ac(FQ1, 1, 0, 0) - ac(FQ1, 0, 0, 0)) * dxi;
            double dudy = (ac(FQ0, 1, 0, 0) - ac(FQ0, 1, -1, 0)) * ac.rw(RDYCI);
            const double t21f = ac(FMS0, 1, 0, 0) * (dvdx + dudy);
            dvdx = (ac(FQ1, 0, 0, 0) - ac(FQ1, -1, 0, 0)) * dxi;
            dudy = (ac(FQ0, 0, 0, 0) - ac(FQ0, 0, -1, 0)) * ac.rw(RDYCI);
            const double t21b = ac(FMS0, 0, 0, 0) * (dvdx + dudy);
            const double d21 = (t21f - t21b) * v1;
            double dvdz = (ac(FQ1, 0, 0, 1) - ac(FQ1, 0, 0, 0)) * dzi;
            double dwdy = (ac(FQ2, 0, 0, 1) - ac(FQ2, 0, -1, 1)) * ac.rw(RDYCI);
            const double t23f = ac(FMS2, 0, 0, 1) * (dvdz + dwdy);
            dvdz = (ac(FQ1, 0, 0, 0) - ac(FQ1, 0, 0, -1)) * dzi;
            dwdy = (ac(FQ2, 0, 0, 0) - ac(FQ2, 0, -1, 0)) * ac.rw(RDYCI);
            const double t23b = ac(FMS2, 0, 0, 0) * (dvdz + dwdy);
            const double d23 = (t23f - t23b) * v3;
            acc = acc + d21 + d22 + d23;
            }
            pgm = (ac(FPR, 0, 0, 0) - ac(FPR, 0, -1, 0)) * (ta * dt * ac.rw(RDYCI));
        } else {
            const double c1 = dxi * hf * hf, c2 = ac.rw(RDYFI) * hf, c3 = dzi * hf * hf;
            double f = (ac(FG0, 1, 0, 0) + ac(FG0, 1, 0, -1)) * (ac(FQ2, 1, 0, 0) + ac(FQ2, 0, 0, 0));
            double b = (ac(FG0, 0, 0, 0) + ac(FG0, 0, 0, -1)) * (ac(FQ2, 0, 0, 0) + ac(FQ2, -1, 0, 0));
            const double h31 = (f - b) * c1;
            f = (ac(FG1, 0, 1, 0) + ac(FG1, 0, 1, -1)) * (wffp * ac(FQ2, 0, 1, 0) + wfbp * ac(FQ2, 0, 0, 0));
            b = (ac(FG1, 0, 0, 0) + ac(FG1, 0, 0, -1)) * (wff * ac(FQ2, 0, 0, 0) + wfb * ac(FQ2, 0, -1, 0));
            const double h32 = (f - b) * c2;
            f = (ac(FG2, 0, 0, 1) + ac(FG2, 0, 0, 0)) * (ac(FQ2, 0, 0, 1) + ac(FQ2, 0, 0, 0));
            b = (ac(FG2, 0, 0, 0) + ac(FG2, 0, 0, -1)) * (ac(FQ2, 0, 0, 0) + ac(FQ2, 0, 0, -1));
            const double h33 = (f - b) * c3;
            acc = -(h31 + h32 + h33);
            if constexpr (CN) {
                const double e1 = dxi * cv, e3 = -2.0 / 3.0 * dzi * cv, e2 = ac.rw(RDYFI) * cv;
                double dudx = (ac(FQ0, 1, 0, 0) - ac(FQ0, 0, 0, 0)) * dxi;
                double dvdy = (ac(FQ1, 0, 1, 0) - ac(FQ1, 0, 0, 0)) * ac.rw(RDYFI);
                double dwdz = (ac(FQ2, 0, 0, 1) - ac(FQ2, 0, 0, 0)) * dzi;
                const double t33fe = ac(FM1, 0, 0, 0) * (dudx + dvdy), t33fi = ac(FM1, 0, 0, 0) * dwdz;
                dudx = (ac(FQ0, 1, 0, -1) - ac(FQ0, 0, 0, -1)) * dxi;
                dvdy = (ac(FQ1, 0, 1, -1) - ac(FQ1, 0, 0, -1)) * ac.rw(RDYFI);
                dwdz = (ac(FQ2, 0, 0, 0) - ac(FQ2, 0, 0, -1)) * dzi;
                const double t33be = ac(FM1, 0, 0, -1) * (dudx + dvdy), t33bi = ac(FM1, 0, 0, -1) * dwdz;
                const double d33e = (t33fe - t33be) * e3, d33i = (t33fi - t33bi) * e3 * (-2.0);
                double vis = ac(FMS1, 1, 0, 0);
                double dwdx = (ac(FQ2, 1, 0, 0) - ac(FQ2, 0, 0, 0)) * dxi;
                double dudz = (ac(FQ0, 1, 0, 0) - ac(FQ0, 1, 0, -1)) * dzi;
                const double t31fe = vis * dudz, t31fi = vis * dwdx;
                vis = ac(FMS1, 0, 0, 0);
                dwdx = (ac(FQ2, 0, 0, 0) - ac(FQ2, -1, 0, 0)) * dxi;
                dudz = (ac(FQ0, 0, 0, 0) - ac(FQ0, 0, 0, -1)) * dzi;
                const double t31be = vis * dudz, t31bi = vis * dwdx;
                const double d31e = (t31fe - t31be) * e1, d31i = (t31fi - t31bi) * e1;
                vis = ac(FMS2, 0, 1, 0);
                double dwdy = (ac(FQ2, 0, 1, 0) - ac(FQ2, 0, 0, 0)) * ac.rw(RDYCIP);
                double dvdz = (ac(FQ1, 0, 1, 0) - ac(FQ1, 0, 1, -1)) * dzi;
                const double t32fe = vis * dvdz, t32fi = vis * dwdy;
                vis = ac(FMS2, 0, 0, 0);
                dwdy = (ac(FQ2, 0, 0, 0) - ac(FQ2, 0, -1, 0)) * ac.rw(RDYCI);
                dvdz = (ac(FQ1, 0, 0, 0) - ac(FQ1, 0, 0, -1)) * dzi;
                const double t32be = vis * dvdz, t32bi = vis * dwdy;
                const double d32e = (t32fe - t32be) * e2, d32i = (t32fi - t32bi) * e2;
                acc = d31e + d32e + d33e + acc;
                imp = d31i + d32i + d33i;
            } else {
            const double v1 = dxi * cv, v2 = ac.rw(RDYFI) * cv, v3 = 2.0 * dzi * cv;
            double dwdz = (ac(FQ2, 0, 0, 1) - ac(FQ2, 0, 0, 0)) * dzi;
            const double t33f = 0.5 * (ac(FM0, 0, 0, 0) + ac(FM0, 0, 0, 0)) * (dwdz - ac(FDU, 0, 0, 0));
            dwdz = (ac(FQ2, 0, 0, 0) - ac(FQ2, 0, 0, -1)) * dzi;
            const double t33b = 0.5 * (ac(FM0, 0, 0, -1) + ac(FM0, 0, 0, -1)) * (dwdz - ac(FDU, 0, 0, -1));
            const double d33 = (t33f - t33b) * v3;
            double dwdx = (ac(FQ2, 1, 0, 0) - ac(FQ2, 0, 0, 0)) * dxi;
            double dudz = (ac(FQ0, 1, 0, 0) - ac(FQ0, 1, 0, -1)) * dzi;
            const double t31f = ac(FMS1, 1, 0, 0) * (dwdx + dudz);
            dwdx = (ac(FQ2, 0, 0, 0) - ac(FQ2, -1, 0, 0)) * dxi;
            dudz = (ac(FQ0, 0, 0, 0) - ac(FQ0, 0, 0, -1)) * dzi;
            const double t31b = ac(FMS1, 0, 0, 0) * (dwdx + dudz);
            const double d31 = (t31f - t31b) * v1;
            double dwdy = (ac(FQ2, 0, 1, 0) - ac(FQ2, 0, 0, 0)) * ac.rw(RDYCIP);
            double dvdz = (ac(FQ1, 0, 1, 0) - ac(FQ1, 0, 1, -1)) * dzi;
            const double t32f = ac(FMS2, 0, 1, 0) * (dwdy + dvdz);
            dwdy = (ac(FQ2, 0, 0, 0) - ac(FQ2, 0, -1, 0)) * ac.rw(RDYCI);
            dvdz = (ac(FQ1, 0, 0, 0) - ac(FQ1, 0, 0, -1)) * dzi;
            const double t32b = ac(FMS2, 0, 0, 0) * (dwdy + dvdz);
            const double d32 = (t32f - t32b) * v2;
            acc = acc + d31 + d32 + d33;
            }
            pgm = (ac(FPR, 0, 0, 0) - ac(FPR, 0, 0, -1)) * (ta * dt * dzi);
        }
        double r;
        if constexpr (CN) r = (tg * acc + tr * ac.ex(C)) * dt + (ta * dt) * imp;
        else r = (tg * acc + tr * ac.ex(C)) * dt;

        r = r - pgm;
        if (d.gravity == C + 1) r = r + ac.ds(C) * (ta * dt * d.f_a);
        ac.put(C, acc, r);
        return C == 0 ? r / ac.rw(RDYFI) : 0.0;
    }
}

// All three momentum components of a point in one pass, so G, Q, the viscosities and the
// pressure are read once for the three; the block's partial of sum(RHS_u / DYFI) for the
// mass-flux correction (RHS_MOM_EXPLICIT_io.f90:229-265) is written as well.
// A 64 x 4 block covers TY rows of TZ = 4 / TY z-planes (TY 4, 2 or 1), so the z-neighbour
// values a point reads belong to the block's own other planes when TY < 4.
template <bool CN, typename W>
__global__ __launch_bounds__(256) void k_io_mom(IoDev d, const double *__restrict__ gs, double tg, double tr, double ta,
                                                double dt, int TY)
{
    __shared__ double sh[256];
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int J = 1 + (int)(blockIdx.y * TY + threadIdx.y % TY);
    const int k = (int)(blockIdx.z * (4 / TY) + threadIdx.y / TY);
    const int L = d.L, M = d.M, N = d.N;
    double contrib = 0.0;
    if (i < L && J <= M && k < N) {
        const IoAtHbm<W> ac{d, gs, i, J, k, i == L - 1 ? 0 : i + 1, i == 0 ? L - 1 : i - 1, k == N - 1 ? 0 : k + 1,
                         k == 0 ? N - 1 : k - 1};
        contrib = io_mom_point<0, CN>(d, ac, i, J, k, tg, tr, ta, dt);
        if (J >= 2) io_mom_point<1, CN>(d, ac, i, J, k, tg, tr, ta, dt);
        io_mom_point<2, CN>(d, ac, i, J, k, tg, tr, ta, dt);
    }
    const double s = block_sum256(contrib, sh);
    if (threadIdx.x == 0 && threadIdx.y == 0)
        d.partial[((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = s;
}

// The same three components from LDS, marching z.  A block of 64 x BY threads owns columns
// i0..i0+63 and rows J0..J0+BY-1 and marches KZ z-planes.  Each field the stencil reads is
// staged per plane as a (BY + 2) x 66 tile (a halo row and column each side, x periodic) in a
// ring of two plane slots, planes k and k+1; the plane-(k-1) values the stencil reads stay in
// registers from the previous plane.  So a value is loaded from HBM once per block and plane
// (halo cells by the neighbouring blocks as well) rather than once per stencil use, and the
// stencil's ~110 loads of a point become LDS reads.  Plane k+2 is loaded into registers while
// plane k is computed and stored into plane k's slot after it.  Identical arithmetic to
// k_io_mom; the block's partial sums its points of all its planes.
template <bool CN> struct IoStage {
    static constexpr int NF = CN ? 11 : 12;   // staged fields: G, Q, MU_STG (3 each), viscosity, p (, div)
    static constexpr int NC = CN ? 12 : 13;   // plane-(k-1) values carried in registers
    // staged slot of field f
    static __device__ __forceinline__ constexpr int slot(int f)
    {
        return f <= FMS2 ? f : (f == FPR ? 10 : (f == FDU ? 11 : 9));
    }
    static __device__ __forceinline__ constexpr int field(int s)
    {
        return s <= FMS2 ? s : (s == 9 ? (CN ? FM1 : FM0) : (s == 10 ? FPR : FDU));
    }
    // register of the carried value (f, di, dj) at plane k-1, -1 if it is not carried
    static __device__ __forceinline__ constexpr int carry(int f, int di, int dj)
    {
        return (f == FQ0 && di == 0 && dj == 0) ? 0 : (f == FQ0 && di == 1 && dj == 0) ? 1
             : (f == FQ1 && di == 0 && dj == 0) ? 2 : (f == FQ1 && di == 0 && dj == 1) ? 3
             : (f == FQ2 && di == 0 && dj == 0) ? 4 : (f == FG0 && di == 0 && dj == 0) ? 5
             : (f == FG0 && di == 1 && dj == 0) ? 6 : (f == FG1 && di == 0 && dj == 0) ? 7
             : (f == FG1 && di == 0 && dj == 1) ? 8 : (f == FG2 && di == 0 && dj == 0) ? 9
             : (f == FPR && di == 0 && dj == 0) ? 10 : (f == (CN ? FM1 : FM0) && di == 0 && dj == 0) ? 11
             : (!CN && f == FDU && di == 0 && dj == 0) ? 12 : -1;
    }
    static __device__ __forceinline__ constexpr int carry_f(int c)
    {
        return c <= 1 ? FQ0 : c <= 3 ? FQ1 : c == 4 ? FQ2 : c <= 6 ? FG0 : c <= 8 ? FG1 : c == 9 ? FG2
             : c == 10 ? FPR : c == 11 ? (CN ? FM1 : FM0) : FDU;
    }
    static __device__ __forceinline__ constexpr int carry_di(int c) { return (c == 1 || c == 6) ? 1 : 0; }
    static __device__ __forceinline__ constexpr int carry_dj(int c) { return (c == 3 || c == 8) ? 1 : 0; }
};

template <bool CN, int BY, typename W> struct IoAtLds {
    static constexpr int TW = IO_TW, TS = (BY + 2) * TW;
    const double *c0, *c1;   // this point's cell in the tiles of planes k and k+1
    const double *cr;        // the carried plane-(k-1) values
    const double *xr;        // EXPLT0 of the three components and the buoyancy component's D_STG
    const double *rr;        // the row's mesh coefficients (IoRowC), loaded once
    IoAtHbm<W> hb;           // any other plane-(k-1) value (the stencil reads none); the stores
    __device__ __forceinline__ double rw(int c) const { return rr[c]; }
    __device__ __forceinline__ double ex(int c) const { return xr[c]; }
    __device__ __forceinline__ double ds(int) const { return xr[3]; }
    __device__ __forceinline__ void put(int c, double acc, double r) const { hb.put(c, acc, r); }
    __device__ __forceinline__ double operator()(int f, int di, int dj, int dk) const
    {
        if (dk < 0) {
            const int c = IoStage<CN>::carry(f, di, dj);
            return c >= 0 ? cr[c] : hb(f, di, dj, dk);
        }
        return (dk > 0 ? c1 : c0)[IoStage<CN>::slot(f) * TS + dj * TW + di];
    }
};

template <bool CN, int BY, typename W>
__global__ __launch_bounds__(64 * BY, 2) void k_io_mom_z(IoDev d, const double *__restrict__ gs, double tg, double tr,
                                                         double ta, double dt, int KZ, int nxcd)
{
    using S = IoStage<CN>;
    constexpr int NT = 64 * BY, TW = IO_TW, TS = (BY + 2) * TW, NF = S::NF, NCH = TS / 2;
    static_assert(NCH <= NT, "one column pair per thread");
    __shared__ __align__(16) double tiles[2][NF * TS];
    __shared__ double sh[NT];
    const int L = d.L, M = d.M, N = d.N;
    const int tx = threadIdx.x, ty = __builtin_amdgcn_readfirstlane(threadIdx.y), t = ty * 64 + tx;
    const int gx = gridDim.x, gy = gridDim.y;
    const int tile = io_xcd_tile(blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z), gx * gy * gridDim.z, nxcd);
    const int bx = tile % gx, by = (tile / gx) % gy, bz = tile / (gx * gy);
    const int i0 = bx * 64, J0 = 1 + by * BY, k0 = bz * KZ, k1 = min(N, k0 + KZ);
    const int i = i0 + tx, J = J0 + ty;
    const bool act = i < L && J <= M;
    const size_t hs = hsz(d), ph = (size_t)(M + 2) * L, pc = (size_t)M * L;
    // in-plane offset of the pair of tile columns this thread stages (pair q of row r: columns
    // i0-2+2q and the next, x periodic; L is even, so a pair never straddles the wrap and
    // starts 16-byte aligned); a row past the field's last (or before a cell field's first)
    // reads the nearest row instead, which no point of the block uses, so that every staging
    // load is unconditional
    const bool stg = t < NCH;
    unsigned oh, oc;
    {
        const int n = stg ? t : 0, r = n / (TW / 2), q = n % (TW / 2), jj = J0 - 1 + r;
        const int ii = ((i0 - 2 + 2 * q) % L + L) % L;
        oh = (unsigned)(min(jj, M + 1) * L + ii);
        oc = (unsigned)((min(max(jj, 1), M) - 1) * L + ii);
    }
    auto base = [&](int f) -> const double * {
        switch (f) {
        case FG0: return gs;
        case FG1: return gs + hs;
        case FG2: return gs + 2 * hs;
        case FQ0: return d.q;
        case FQ1: return d.q + hs;
        case FQ2: return d.q + 2 * hs;
        case FMS0: return d.mstg;
        case FMS1: return d.mstg + hs;
        case FMS2: return d.mstg + 2 * hs;
        case FM0: return d.m0;
        case FM1: return d.m1;
        case FPR: return d.pr;
        default: return d.divu;
        }
    };
    double pre0[NF], pre1[NF];   // the staged pair (two arrays: a double2 array went to scratch)
    auto load = [&](int kk) {
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            const int fld = S::field(f);
            const bool cell = fld == FDU;
            const double *b = base(fld) + (size_t)kk * (cell ? pc : ph);
            const double2 v = *reinterpret_cast<const double2 *>(b + (cell ? oc : oh));
            pre0[f] = v.x;
            pre1[f] = v.y;
        }
    };
    auto store = [&](int sl) {
        if (stg) {
#pragma unroll
            for (int f = 0; f < NF; ++f)
                reinterpret_cast<double2 *>(&tiles[sl][f * TS])[t] = make_double2(pre0[f], pre1[f]);
        }
    };
    const int ip = i == L - 1 ? 0 : i + 1, im = i == 0 ? L - 1 : i - 1;
    double cr[S::NC];
    {
        const int km = k0 == 0 ? N - 1 : k0 - 1;
        const IoAtHbm<> hb{d, gs, act ? i : 0, act ? J : 1, km, act ? ip : 0, act ? im : 0, km, km};
#pragma unroll
        for (int c = 0; c < S::NC; ++c) cr[c] = hb(S::carry_f(c), S::carry_di(c), S::carry_dj(c), 0);
    }
    // EXPLT0 of the three components and the buoyancy component's D_STG at this point, loaded
    // a plane ahead and before that plane's tiles: the vector-memory counter retires loads in
    // issue order, so a load issued after the tile loads would make the plane wait for them
    const int gc = d.gravity - 1;
    double xr[4] = {0.0, 0.0, 0.0, 0.0}, xn[4] = {0.0, 0.0, 0.0, 0.0};
    auto loadx = [&](int kk, double (&x)[4]) {
        if (act) {
            const size_t o = cx(d, i, J, kk);
#pragma unroll
            for (int c = 0; c < 3; ++c) x[c] = ioex<W>(d)[(size_t)c * pc * N + o];
            if (gc >= 0 && gc <= 2) x[3] = d.dstg[(size_t)gc * hs + hx(d, i, J, kk)];
        }
    };
    double rr[IO_NROW];
    {
        const int Jr = min(J, M);   // rows past M (no points) take row M's
        const IoAtHbm<> hb{d, gs, 0, Jr, 0, 0, 0, 0, 0};
#pragma unroll
        for (int c = 0; c < IO_NROW; ++c) rr[c] = (c == RDYFIM && Jr < 2) ? 0.0 : hb.rw(c);
    }
    loadx(k0, xr);
    load(k0);
    store(0);
    load(k0 + 1 == N ? 0 : k0 + 1);
    store(1);
    __syncthreads();
    const int cell = (ty + 1) * TW + tx + 2;
    double contrib = 0.0;
    for (int k = k0; k < k1; ++k) {
        const int cur = (k - k0) & 1;
        const bool more = k + 1 < k1;
        if (more) {
            loadx(k + 1, xn);
            load(k + 2 >= N ? k + 2 - N : k + 2);
        }
        if (act) {
            const IoAtLds<CN, BY, W> ac{&tiles[cur][cell], &tiles[cur ^ 1][cell], cr, xr, rr,
                                        IoAtHbm<W>{d, gs, i, J, k, ip, im, k == N - 1 ? 0 : k + 1, k == 0 ? N - 1 : k - 1}};
            contrib = contrib + io_mom_point<0, CN>(d, ac, i, J, k, tg, tr, ta, dt);
            if (J >= 2) io_mom_point<1, CN>(d, ac, i, J, k, tg, tr, ta, dt);
            io_mom_point<2, CN>(d, ac, i, J, k, tg, tr, ta, dt);
        }
#pragma unroll
        for (int c = 0; c < S::NC; ++c)
            cr[c] = tiles[cur][S::slot(S::carry_f(c)) * TS + cell + S::carry_dj(c) * TW + S::carry_di(c)];
#pragma unroll
        for (int c = 0; c < 4; ++c) xr[c] = xn[c];
        __syncthreads();
        if (more) store(cur);
        __syncthreads();
    }
    sh[t] = contrib;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] = sh[t] + sh[t + s];
        __syncthreads();
    }
    if (t == 0) d.partial[tile] = sh[0];
}

// DPGRNS = sum of the partials / VOLM_io, one block, fixed order.
__global__ __launch_bounds__(256) void k_io_reduce(const double *__restrict__ part, int n, double volm,
                                                   double *__restrict__ out)
{
    __shared__ double sh[256];
    double v = 0.0;
    for (int j = threadIdx.x; j < n; j += 256) v = v + part[j];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0] / volm;
}

// The streamwise correction (RHS_MOM_EXPLICIT_io.f90:259-265) and MASSFLUX_CALC_io
// (MASSFLUX_CALC_IO.f90:15-52) with BC_WALL_G_io: rows 0..M+1.
// gs: G at the start of the substep (G0_io); writes every row of the new G.
template <typename W>
__global__ __launch_bounds__(256) void k_io_update_g(IoDev d, const double *__restrict__ gs,
                                                     const double *__restrict__ dpg_dev, double dpg)
{
#pragma clang fp contract(off)
    IO_THREAD(0)
    if (i >= L || J > M + 1) return;
    const size_t hs = hsz(d), c0 = hx(d, i, J, k);
    if (J == 0 || J == M + 1) {
        for (int c = 0; c < 3; ++c) d.g[c * hs + c0] = 0.0;
        return;
    }
    const double dp = dpg_dev ? dpg_dev[0] : dpg;
    const size_t ci = cx(d, i, J, k), cs = csz(d);
    const W *h = ioh<W>(d);
    d.g[c0] = gs[c0] + ((double)h[ci] - dp);
    d.g[2 * hs + c0] = gs[2 * hs + c0] + (double)h[2 * cs + ci];
    d.g[hs + c0] = J == 1 ? 0.0 : gs[hs + c0] + (double)h[cs + ci];
}

// ---- Crank-Nicolson line solves (MOMFA_io, MOMFA_io.f90:18-406) --------------------------
// One thread per line, each following the reference's Thomas recurrences in their order
// (TDMAIJI_CYC.f90:88-157, TDMAIJJ_nonCYC.f90:88-123) with the coefficients built per element
// as MOMFA1_X_io / MOMFA2_Z_io / MOMFA3_Y_io build them.  The recurrences' intermediate rows
// go to d.work element-major ([element][line]), so a wave's accesses are contiguous; the three
// momentum components run side by side (grid.y), each on its own third of d.work.

// Cyclic recurrence of TDMAIJI_CYC on one line of n >= 3 elements; coef(i, a, b, c) gives
// row i, rhs(i) its right-hand side, out(i, x) takes the solution.  H, G1, G2: element i at
// [i * es].
template <class Coef, class Rhs, class Out>
__device__ __forceinline__ void thomas_cyc(int n, const Coef &coef, const Rhs &rhs, const Out &out,
                                           double *__restrict__ H, double *__restrict__ G1,
                                           double *__restrict__ G2, size_t es)
{
#pragma clang fp contract(off)
    double a, b, c;
    coef(0, a, b, c);
    double h = c / b, g1 = rhs(0) / b, g2 = -a / b;
    H[0] = h;
    G1[0] = g1;
    G2[0] = g2;
    for (int i = 1; i <= n - 3; ++i) {
        coef(i, a, b, c);
        const double pp = 1.0 / (b - h * a);
        const double qq1 = g1 * a, qq2 = g2 * a;
        h = c * pp;
        g1 = (rhs(i) - qq1) * pp;
        g2 = -qq2 * pp;
        H[i * es] = h;
        G1[i * es] = g1;
        G2[i * es] = g2;
    }
    {
        const int i = n - 2;
        coef(i, a, b, c);
        const double pp = 1.0 / (b - h * a);
        const double qq1 = g1 * a, qq2 = g2 * a;
        g1 = (rhs(i) - qq1) * pp;
        g2 = (-c - qq2) * pp;
        G1[i * es] = g1;
        G2[i * es] = g2;
    }
    const double ge1 = g1, ge2 = g2;   // rows n-2, final
    for (int i = n - 3; i >= 0; --i) {
        const double hi = H[i * es];
        g1 = -hi * g1 + G1[i * es];
        g2 = -hi * g2 + G2[i * es];
        G1[i * es] = g1;
        G2[i * es] = g2;
    }
    const int e = n - 1;
    coef(e, a, b, c);
    const double pp = rhs(e) - c * g1 - a * ge1;
    const double qq1 = b + c * g2 + a * ge2;
    const double xe = pp / qq1;
    out(e, xe);
    out(n - 2, ge1 + ge2 * xe);
    for (int i = n - 3; i >= 0; --i) out(i, G1[i * es] + G2[i * es] * xe);
}

// Recurrence of TDMAIJJ_nonCYC with zero boundary values on one line of n >= 2 elements.
template <class Coef, class Rhs, class Out>
__device__ __forceinline__ void thomas_line(int n, const Coef &coef, const Rhs &rhs, const Out &out,
                                            double *__restrict__ H, double *__restrict__ G, size_t es)
{
#pragma clang fp contract(off)
    const double bc = 0.0;
    double a, b, c;
    coef(0, a, b, c);
    double h = c / b, g = (rhs(0) - a * bc) / b;
    H[0] = h;
    G[0] = g;
    for (int j = 1; j <= n - 2; ++j) {
        coef(j, a, b, c);
        const double pp = 1.0 / (b - h * a);
        h = c * pp;
        g = (rhs(j) - a * g) * pp;
        H[j * es] = h;
        G[j * es] = g;
    }
    coef(n - 1, a, b, c);
    const double pp = 1.0 / (b - h * a);
    double x = (rhs(n - 1) - c * bc - a * g) * pp;
    out(n - 1, x);
    for (int j = n - 2; j >= 0; --j) {
        x = G[j * es] - H[j * es] * x;
        out(j, x);
    }
}

// MOMFA1_X_io (MOMFA_io.f90:77-176): cyclic x-lines of component C over rows J0..M (2..M for v)
// with FACOE = alpha dt / (2 Re); the mass-flux correction of the streamwise increment
// (RHS_MOM_EXPLICIT_io.f90:259-265) is applied as the line is read.  In place on d.h[C].
template <int C, typename W>
__device__ __forceinline__ void io_cn_x_line(const IoDev &d, int l, double dp, double facoe)
{
#pragma clang fp contract(off)
    const int L = d.L, M = d.M, j0 = C == 1 ? 2 : 1, nj = M - j0 + 1, nl = nj * d.N;
    if (l >= nl) return;
    const int J = j0 + l % nj, k = l / nj;
    const size_t hs = hsz(d), cs = csz(d);
    const double *DS = d.dstg + C * hs;
    W *R = ioh<W>(d) + C * cs;
    double *Wk = d.work + (size_t)C * 3 * cs + l;
    const double coe1 = facoe * 4.0 / 3.0 * d.dxqi, coe2 = facoe * d.dxqi;
    auto coef = [&](int i, double &a, double &b, double &c) {
        const int ip = i == L - 1 ? 0 : i + 1, im = i == 0 ? L - 1 : i - 1;
        double amu, apu;
        if constexpr (C == 0) {
            amu = -coe1 * d.m1[hx(d, im, J, k)];
            apu = -coe1 * d.m1[hx(d, i, J, k)];
        } else if constexpr (C == 1) {
            amu = -coe2 * d.mstg[hx(d, i, J, k)];
            apu = -coe2 * d.mstg[hx(d, ip, J, k)];
        } else {
            amu = -coe2 * d.mstg[hs + hx(d, i, J, k)];
            apu = -coe2 * d.mstg[hs + hx(d, ip, J, k)];
        }
        a = amu / DS[hx(d, im, J, k)];
        c = apu / DS[hx(d, ip, J, k)];
        b = 1.0 - (amu + apu) / DS[hx(d, i, J, k)];
    };
    auto rhs = [&](int i) {
        if constexpr (C == 0) return (double)R[cx(d, i, J, k)] - dp;
        else return (double)R[cx(d, i, J, k)];
    };
    auto out = [&](int i, double x) { R[cx(d, i, J, k)] = (W)x; };
    thomas_cyc(L, coef, rhs, out, Wk, Wk + cs, Wk + 2 * cs, (size_t)nl);
}

template <typename W>
__global__ __launch_bounds__(128) void k_io_cn_x(IoDev d, const double *__restrict__ dpg_dev, double dpg, double facoe)
{
    const int l = blockIdx.x * 128 + threadIdx.x;
    if (blockIdx.y == 0) io_cn_x_line<0, W>(d, l, dpg_dev ? dpg_dev[0] : dpg, facoe);
    else if (blockIdx.y == 1) io_cn_x_line<1, W>(d, l, 0.0, facoe);
    else io_cn_x_line<2, W>(d, l, 0.0, facoe);
}

// Block maximum of a non-negative value: a wave maximum by shuffles, one 64-bit atomic per
// wave on an LDS word (for non-negative doubles the bit order is the value order).  Every
// thread of the block must call it; blockDim.x must be a multiple of 64.
__device__ __forceinline__ double block_max_nonneg(double v, unsigned long long *word)
{
    if (threadIdx.x == 0) *word = 0ull;
    __syncthreads();
    for (int o = 32; o > 0; o >>= 1) v = nanmax(v, __shfl_xor(v, o));
    if ((threadIdx.x & 63) == 0) atomicMax(word, (unsigned long long)__double_as_longlong(v));
    __syncthreads();
    return __longlong_as_double((long long)*word);
}

// The Crank-Nicolson factors are strictly diagonally dominant with a small ratio
// rho = max (|a| + |c|) / |b| (about 1e-4 in x and z and 2e-3 at the walls in y on the
// channel meshes).  Starting from x0 = r / b, K Jacobi sweeps x <- (r - a x_- - c x_+) / b leave
// an error of at most rho^(K+1) max|x| (|x0 - x| <= rho |x|, each sweep contracts by rho); K is
// the least count with rho^(K+1) <= 1e-17, e.g. 3 sweeps at rho = 1e-4.  The line kernels below
// solve this way and hand a block whose rho exceeds 1/4 to the thread-per-line recurrences
// (io_cn_x_line, io_cn_z_line, io_cn_y_line).
// 1/x from the hardware estimate and two Newton-Raphson steps r <- r + r (1 - x r), each of
// which doubles the correct bits: within about an ulp of the correctly rounded reciprocal.
__device__ __forceinline__ double rcp_nr(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    r = fma(fma(-x, r, 1.0), r, r);
    return fma(fma(-x, r, 1.0), r, r);
}

// a / b for the Jacobi line solves' coefficients: the IEEE quotient, or with FD a times
// rcp_nr(b) (one reciprocal per density value, shared by the neighbouring rows' coefficients).
// The sweeps approximate the exact solve to 1e-17 relative, so an ulp in a coefficient is
// below what they resolve.
template <bool FD> __device__ __forceinline__ double qd(double a, double b) { return FD ? a * rcp_nr(b) : a / b; }

__device__ __forceinline__ int jacobi_sweeps(double rho)
{
    int K = 0;
    for (double q = rho; q > 1e-17 && K < 60; q *= rho) ++K;
    return K;
}

// MOMFA1_X_io for L a power of two (4 <= L <= 512): T = blockDim.x / L x-lines per block, one
// thread per point, coalesced; the coefficients are built from the staged property rows as
// io_cn_x_line builds them.  Lines past the end carry a dummy identity row.
template <int C, typename W>
__device__ __forceinline__ void io_cn_x_jac(const IoDev &d, double dp, double facoe, double *lds)
{
#pragma clang fp contract(off)
    const int L = d.L, M = d.M, j0 = C == 1 ? 2 : 1, nj = M - j0 + 1, nl = nj * d.N;
    const int T = blockDim.x / L, lt = threadIdx.x / L, e = threadIdx.x % L;
    const int line = blockIdx.x * T + lt;
    const bool valid = line < nl;
    const int J = j0 + line % nj, k = line / nj;
    const size_t hs = hsz(d), cs = csz(d);
    const int S = blockDim.x;
    double *s0 = lds, *s1 = lds + S;
    unsigned long long *word = reinterpret_cast<unsigned long long *>(lds + 2 * S);
    const int sm = lt * L + ((e - 1) & (L - 1)), sp = lt * L + ((e + 1) & (L - 1));
    double f0 = 0.0, f1 = 1.0, r = 0.0;
    if (valid) {
        const size_t o = hx(d, e, J, k);
        f0 = C == 0 ? d.m1[o] : d.mstg[(C - 1) * hs + o];
        f1 = d.dstg[C * hs + o];
        r = ioh<W>(d)[C * cs + cx(d, e, J, k)];
        if constexpr (C == 0) r = r - dp;
    }
    s0[threadIdx.x] = f0;
    s1[threadIdx.x] = f1;
    __syncthreads();
    double amu, apu;
    if constexpr (C == 0) {
        const double coe1 = facoe * 4.0 / 3.0 * d.dxqi;
        amu = -coe1 * s0[sm];
        apu = -coe1 * f0;
    } else {
        const double coe2 = facoe * d.dxqi;
        amu = -coe2 * f0;
        apu = -coe2 * s0[sp];
    }
    double a = amu / s1[sm], c = apu / s1[sp], b = 1.0 - (amu + apu) / f1;
    if (!valid) {
        a = c = 0.0;
        b = 1.0;
    }
    const double ib = 1.0 / b;
    const double rho = block_max_nonneg((fabs(a) + fabs(c)) * fabs(ib), word);
    if (!(rho <= 0.25)) {   // the same branch for the whole block
        if (valid && e == 0) io_cn_x_line<C, W>(d, line, dp, facoe);
        return;
    }
    const int K = jacobi_sweeps(rho);
    double x = r * ib;
    for (int it = 0; it < K; ++it) {
        double *sb = (it & 1) ? s1 : s0;   // alternating buffers: one barrier per sweep
        sb[threadIdx.x] = x;
        __syncthreads();
        const double xm = sb[sm], xp = sb[sp];
        x = (r - a * xm - c * xp) * ib;
    }
    if (valid) ioh<W>(d)[C * cs + cx(d, e, J, k)] = (W)x;
}

// MOMFA1_X_io for L = 64 P (P = 1..8): one wavefront per x-line, lane l holding points
// l P .. l P + P - 1; the neighbours across lanes come by cross-lane shuffles, so a line needs
// no block barrier and a wavefront's loads are all issued together.  Coefficients and sweeps
// as io_cn_x_jac (rho is the line's own maximum).
template <int C, int P, typename W, bool FD>
__device__ __forceinline__ void io_cn_x_wave(const IoDev &d, double dp, double facoe)
{
#pragma clang fp contract(off)
    const int M = d.M, j0 = C == 1 ? 2 : 1, nj = M - j0 + 1, nl = nj * d.N;
    const int lane = threadIdx.x & 63;
    const int line = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    if (line >= nl) return;   // the whole wavefront
    const int J = j0 + line % nj, k = line / nj;
    const size_t hs = hsz(d), cs = csz(d);
    const int lm = (lane + 63) & 63, lp = (lane + 1) & 63;
    const double *F0 = (C == 0 ? d.m1 : d.mstg + (C - 1) * hs) + hx(d, lane * P, J, k);
    const double *F1 = d.dstg + C * hs + hx(d, lane * P, J, k);
    W *R = ioh<W>(d) + C * cs + cx(d, lane * P, J, k);
    double f0[P], f1[P], r[P];
#pragma unroll
    for (int j = 0; j < P; ++j) {
        f0[j] = F0[j];
        f1[j] = F1[j];
        r[j] = R[j];
        if constexpr (C == 0) r[j] = r[j] - dp;
    }
    const double f0m = __shfl(f0[P - 1], lm), f0p = __shfl(f0[0], lp);
    const double f1m = __shfl(f1[P - 1], lm), f1p = __shfl(f1[0], lp);
    double a[P], c[P], ib[P], rho = 0.0;
#pragma unroll
    for (int j = 0; j < P; ++j) {
        const double vm = j > 0 ? f0[j - 1] : f0m, vp = j < P - 1 ? f0[j + 1] : f0p;
        const double dm = j > 0 ? f1[j - 1] : f1m, dq = j < P - 1 ? f1[j + 1] : f1p;
        double amu, apu;
        if constexpr (C == 0) {
            const double coe1 = facoe * 4.0 / 3.0 * d.dxqi;
            amu = -coe1 * vm;
            apu = -coe1 * f0[j];
        } else {
            const double coe2 = facoe * d.dxqi;
            amu = -coe2 * f0[j];
            apu = -coe2 * vp;
        }
        a[j] = qd<FD>(amu, dm);
        c[j] = qd<FD>(apu, dq);
        const double b = 1.0 - qd<FD>(amu + apu, f1[j]);
        ib[j] = qd<FD>(1.0, b);
        rho = nanmax(rho, (fabs(a[j]) + fabs(c[j])) * fabs(ib[j]));
    }
    for (int o = 32; o > 0; o >>= 1) rho = nanmax(rho, __shfl_xor(rho, o));
    if (!(rho <= 0.25)) {
        if (lane == 0) io_cn_x_line<C, W>(d, line, dp, facoe);
        return;
    }
    const int K = jacobi_sweeps(rho);
    double x[P];
#pragma unroll
    for (int j = 0; j < P; ++j) x[j] = r[j] * ib[j];
    for (int it = 0; it < K; ++it) {
        const double xl = __shfl(x[P - 1], lm), xr = __shfl(x[0], lp);
        double xn[P];
#pragma unroll
        for (int j = 0; j < P; ++j) {
            const double xm = j > 0 ? x[j - 1] : xl, xp = j < P - 1 ? x[j + 1] : xr;
            xn[j] = (r[j] - a[j] * xm - c[j] * xp) * ib[j];
        }
#pragma unroll
        for (int j = 0; j < P; ++j) x[j] = xn[j];
    }
#pragma unroll
    for (int j = 0; j < P; ++j) R[j] = (W)x[j];
}

template <int P, typename W, bool FD>
__global__ __launch_bounds__(256) void k_io_cn_x_wave(IoDev d, const double *__restrict__ dpg_dev, double dpg,
                                                      double facoe)
{
    if (blockIdx.y == 0) io_cn_x_wave<0, P, W, FD>(d, dpg_dev ? dpg_dev[0] : dpg, facoe);
    else if (blockIdx.y == 1) io_cn_x_wave<1, P, W, FD>(d, 0.0, facoe);
    else io_cn_x_wave<2, P, W, FD>(d, 0.0, facoe);
}

template <typename W>
__global__ __launch_bounds__(512) void k_io_cn_x_jac(IoDev d, const double *__restrict__ dpg_dev, double dpg,
                                                     double facoe)
{
    extern __shared__ double lds[];
    if (blockIdx.y == 0) io_cn_x_jac<0, W>(d, dpg_dev ? dpg_dev[0] : dpg, facoe, lds);
    else if (blockIdx.y == 1) io_cn_x_jac<1, W>(d, 0.0, facoe, lds);
    else io_cn_x_jac<2, W>(d, 0.0, facoe, lds);
}

// MOMFA2_Z_io (MOMFA_io.f90:179-265): cyclic z-lines (RCCI2 = RNDI2 = 1 on the channel).
template <int C, typename W>
__device__ __forceinline__ void io_cn_z_line(const IoDev &d, int l, double facoe)
{
#pragma clang fp contract(off)
    const int L = d.L, M = d.M, N = d.N, j0 = C == 1 ? 2 : 1, nl = (M - j0 + 1) * L;
    if (l >= nl) return;
    const int i = l % L, J = j0 + l / L;
    const size_t hs = hsz(d), cs = csz(d);
    const double *DS = d.dstg + C * hs;
    W *R = ioh<W>(d) + C * cs;
    double *Wk = d.work + (size_t)C * 3 * cs + l;
    const double coe1 = facoe * 4.0 / 3.0 * d.dzqi, coe2 = facoe * d.dzqi;
    auto coef = [&](int k, double &a, double &b, double &c) {
        const int kp = k == N - 1 ? 0 : k + 1, km = k == 0 ? N - 1 : k - 1;
        double amu, apu;
        if constexpr (C == 0) {
            amu = -coe2 * d.mstg[hs + hx(d, i, J, k)];
            apu = -coe2 * d.mstg[hs + hx(d, i, J, kp)];
        } else if constexpr (C == 1) {
            amu = -coe2 * d.mstg[2 * hs + hx(d, i, J, k)];
            apu = -coe2 * d.mstg[2 * hs + hx(d, i, J, kp)];
        } else {
            amu = -coe1 * d.m1[hx(d, i, J, km)];
            apu = -coe1 * d.m1[hx(d, i, J, k)];
        }
        a = amu / DS[hx(d, i, J, km)];
        c = apu / DS[hx(d, i, J, kp)];
        b = 1.0 - (amu + apu) / DS[hx(d, i, J, k)];
    };
    auto rhs = [&](int k) { return (double)R[cx(d, i, J, k)]; };
    auto out = [&](int k, double x) { R[cx(d, i, J, k)] = (W)x; };
    thomas_cyc(N, coef, rhs, out, Wk, Wk + cs, Wk + 2 * cs, (size_t)nl);
}

template <typename W>
__global__ __launch_bounds__(128) void k_io_cn_z(IoDev d, double facoe)
{
    const int l = blockIdx.x * 128 + threadIdx.x;
    if (blockIdx.y == 0) io_cn_z_line<0, W>(d, l, facoe);
    else if (blockIdx.y == 1) io_cn_z_line<1, W>(d, l, facoe);
    else io_cn_z_line<2, W>(d, l, facoe);
}

// MOMFA3_Y_io (MOMFA_io.f90:268-406): y-lines over rows 1..M with zero boundary values, then
// G = increment + G0 for rows N2I..M (MOMFA_io.f90:38-44) with BC_WALL_G_io: every row of the
// line's new G.  u, w: the wall rows' viscosity is the mean of the wall cell and its ghost in
// the line's own z-plane (the one-rank form of MOMFA_io.f90:342-352).  v: row 1 is the wall,
// AMVR = APVR = 0 there, RHS 0.
template <int C, typename W>
__device__ __forceinline__ void io_cn_y_line(const IoDev &d, const double *__restrict__ gs, int l, double facoe)
{
#pragma clang fp contract(off)
    const int L = d.L, M = d.M, nl = L * d.N;
    if (l >= nl) return;
    const int i = l % L, k = l / L;
    const size_t hs = hsz(d), cs = csz(d);
    const double *DS = d.dstg + C * hs, *FM = C == 1 ? d.m1 : d.mstg + C * hs;
    const double *am = d.amvr + C * (M + 2), *ap = d.apvr + C * (M + 2);
    const W *R = ioh<W>(d) + C * cs;
    double *Wk = d.work + (size_t)C * 3 * cs + l;
    const double coe1 = facoe * 4.0 / 3.0, coe2 = facoe;
    auto coef = [&](int jj, double &a, double &b, double &c) {
        const int J = jj + 1;
        double amu, apu;
        if constexpr (C != 1) {
            if (J == 1) {
                const double muqt = (d.m1[hx(d, i, 1, k)] + d.m1[hx(d, i, 0, k)]) * 0.5;
                amu = -coe2 * d.amvr1 * muqt;
                apu = -coe2 * ap[J] * FM[hx(d, i, J + 1, k)];
                a = 0.0;
                c = apu / DS[hx(d, i, J + 1, k)];
            } else if (J == M) {
                const double muqt = (d.m1[hx(d, i, M, k)] + d.m1[hx(d, i, M + 1, k)]) * 0.5;
                amu = -coe2 * am[J] * FM[hx(d, i, J, k)];
                apu = -coe2 * d.apvrn * muqt;
                a = amu / DS[hx(d, i, J - 1, k)];
                c = 0.0;
            } else {
                amu = -coe2 * am[J] * FM[hx(d, i, J, k)];
                apu = -coe2 * ap[J] * FM[hx(d, i, J + 1, k)];
                a = amu / DS[hx(d, i, J - 1, k)];
                c = apu / DS[hx(d, i, J + 1, k)];
            }
        } else {
            amu = -coe1 * am[J] * FM[hx(d, i, J - 1, k)];
            apu = -coe1 * ap[J] * FM[hx(d, i, J, k)];
            a = J == 1 ? 0.0 : amu / DS[hx(d, i, J - 1, k)];
            c = apu / DS[hx(d, i, J == 1 ? J + 1 : J, k)];
        }
        b = 1.0 - (amu + apu) / DS[hx(d, i, J, k)];
    };
    auto rhs = [&](int jj) { return (C == 1 && jj == 0) ? 0.0 : (double)R[cx(d, i, jj + 1, k)]; };
    double *G = d.g + C * hs;
    const double *G0 = gs + C * hs;
    auto out = [&](int jj, double x) {
        const int J = jj + 1;
        const size_t o = hx(d, i, J, k);
        G[o] = (C == 1 && J == 1) ? 0.0 : x + G0[o];
    };
    G[hx(d, i, 0, k)] = 0.0;
    G[hx(d, i, M + 1, k)] = 0.0;
    thomas_line(M, coef, rhs, out, Wk, Wk + cs, (size_t)nl);
}

template <typename W>
__global__ __launch_bounds__(128) void k_io_cn_y(IoDev d, const double *__restrict__ gs, double facoe)
{
    const int l = blockIdx.x * 128 + threadIdx.x;
    if (blockIdx.y == 0) io_cn_y_line<0, W>(d, gs, l, facoe);
    else if (blockIdx.y == 1) io_cn_y_line<1, W>(d, gs, l, facoe);
    else io_cn_y_line<2, W>(d, gs, l, facoe);
}

// MOMFA2_Z_io with P consecutive z-points per thread (N a multiple of P): a block holds T
// x-columns of one row J, thread (lt, chunk) at lt + T chunk, so a wave reads runs of T
// consecutive x; only the chunk edges go through LDS (coefficient neighbours, then the Jacobi
// sweeps' neighbours).
template <int C, int P, typename W, bool FD>
__device__ __forceinline__ void io_cn_z_jac(const IoDev &d, double facoe, double *lds)
{
#pragma clang fp contract(off)
    const int L = d.L, M = d.M, N = d.N, j0 = C == 1 ? 2 : 1;
    const int NT = N / P, T = blockDim.x / NT, lt = threadIdx.x % T, tt = threadIdx.x / T;
    const int ng = (L + T - 1) / T;
    const int i = (blockIdx.x % ng) * T + lt, J = j0 + blockIdx.x / ng;
    const bool valid = i < L && J <= M;
    const size_t hs = hsz(d), cs = csz(d);
    const int S = blockDim.x;
    double *e0 = lds, *e1 = lds + S, *e2 = lds + 2 * S, *e3 = lds + 3 * S;
    unsigned long long *word = reinterpret_cast<unsigned long long *>(lds + 4 * S);
    const int k0 = tt * P;
    const int sl = ((tt + NT - 1) % NT) * T + lt, sr = ((tt + 1) % NT) * T + lt;   // neighbour chunks
    double f0[P], f1[P], r[P];
#pragma unroll
    for (int j = 0; j < P; ++j) {
        f0[j] = 0.0;
        f1[j] = 1.0;
        r[j] = 0.0;
        if (valid) {
            const size_t o = hx(d, i, J, k0 + j);
            f0[j] = C == 2 ? d.m1[o] : d.mstg[(C + 1) * hs + o];
            f1[j] = d.dstg[C * hs + o];
            r[j] = ioh<W>(d)[C * cs + cx(d, i, J, k0 + j)];
        }
    }
    e0[threadIdx.x] = f0[0];
    e1[threadIdx.x] = f0[P - 1];
    e2[threadIdx.x] = f1[0];
    e3[threadIdx.x] = f1[P - 1];
    __syncthreads();
    const double f0m = e1[sl], f1m = e3[sl], f0p = e0[sr], f1p = e2[sr];
    double a[P], c[P], ib[P], rho = 0.0;
#pragma unroll
    for (int j = 0; j < P; ++j) {
        const double vm = j > 0 ? f0[j - 1] : f0m, vp = j < P - 1 ? f0[j + 1] : f0p;
        const double dm = j > 0 ? f1[j - 1] : f1m, dq = j < P - 1 ? f1[j + 1] : f1p;
        double amu, apu;
        if constexpr (C == 2) {
            const double coe1 = facoe * 4.0 / 3.0 * d.dzqi;
            amu = -coe1 * vm;
            apu = -coe1 * f0[j];
        } else {
            const double coe2 = facoe * d.dzqi;
            amu = -coe2 * f0[j];
            apu = -coe2 * vp;
        }
        a[j] = qd<FD>(amu, dm);
        c[j] = qd<FD>(apu, dq);
        double b = 1.0 - qd<FD>(amu + apu, f1[j]);
        if (!valid) {
            a[j] = c[j] = 0.0;
            b = 1.0;
        }
        ib[j] = qd<FD>(1.0, b);
        rho = nanmax(rho, (fabs(a[j]) + fabs(c[j])) * fabs(ib[j]));
    }
    rho = block_max_nonneg(rho, word);   // also orders the edge reads above before the sweeps
    if (!(rho <= 0.25)) {
        if (valid && tt == 0) io_cn_z_line<C, W>(d, (J - j0) * L + i, facoe);
        return;
    }
    const int K = jacobi_sweeps(rho);
    double x[P];
#pragma unroll
    for (int j = 0; j < P; ++j) x[j] = r[j] * ib[j];
    for (int it = 0; it < K; ++it) {
        // alternating edge buffers: a sweep's writes never meet the previous sweep's reads, so
        // one barrier per sweep
        double *w0 = (it & 1) ? e2 : e0, *w1 = (it & 1) ? e3 : e1;
        w0[threadIdx.x] = x[0];
        w1[threadIdx.x] = x[P - 1];
        __syncthreads();
        const double xl = w1[sl], xr = w0[sr];
        double xn[P];
#pragma unroll
        for (int j = 0; j < P; ++j) {
            const double xm = j > 0 ? x[j - 1] : xl, xp = j < P - 1 ? x[j + 1] : xr;
            xn[j] = (r[j] - a[j] * xm - c[j] * xp) * ib[j];
        }
#pragma unroll
        for (int j = 0; j < P; ++j) x[j] = xn[j];
    }
    if (valid) {
#pragma unroll
        for (int j = 0; j < P; ++j) ioh<W>(d)[C * cs + cx(d, i, J, k0 + j)] = (W)x[j];
    }
}

template <typename W, bool FD, int P>
__global__ __launch_bounds__(1024) void k_io_cn_z_jac(IoDev d, double facoe)
{
    extern __shared__ double lds[];
    if (blockIdx.y == 0) io_cn_z_jac<0, P, W, FD>(d, facoe, lds);
    else if (blockIdx.y == 1) io_cn_z_jac<1, P, W, FD>(d, facoe, lds);
    else io_cn_z_jac<2, P, W, FD>(d, facoe, lds);
}

// MOMFA3_Y_io + the G update with P consecutive rows per thread (NT = ceil(M / P) chunks per
// column, rows past M are identity rows): a block holds T x-columns of one z-plane, thread
// (lt, chunk) at lt + T chunk; coefficients as io_cn_y_line builds them; writes every row of
// the new G of its columns.
template <int C, int P, typename W, bool FD>
__device__ __forceinline__ void io_cn_y_jac(const IoDev &d, const double *__restrict__ gs, double facoe, double *lds)
{
#pragma clang fp contract(off)
    const int L = d.L, M = d.M;
    const int NT = (M + P - 1) / P, T = blockDim.x / NT, lt = threadIdx.x % T, tt = threadIdx.x / T;
    const int ng = (L + T - 1) / T;
    const int i = (blockIdx.x % ng) * T + lt, k = blockIdx.x / ng;
    const bool valid = i < L;
    const size_t hs = hsz(d), cs = csz(d);
    const int S = blockDim.x;
    double *e0 = lds, *e1 = lds + S, *e2 = lds + 2 * S, *e3 = lds + 3 * S;
    unsigned long long *word = reinterpret_cast<unsigned long long *>(lds + 4 * S);
    const int J0 = tt * P + 1;
    const int sl = tt > 0 ? (tt - 1) * T + lt : -1, sr = tt < NT - 1 ? (tt + 1) * T + lt : -1;
    const double *DS = d.dstg + C * hs, *FM = C == 1 ? d.m1 : d.mstg + C * hs;
    double f0[P], f1[P], r[P];
#pragma unroll
    for (int j = 0; j < P; ++j) {
        const int J = J0 + j;
        f0[j] = 0.0;
        f1[j] = 1.0;
        r[j] = 0.0;
        if (valid && J <= M) {
            const size_t o = hx(d, i, J, k);
            f0[j] = FM[o];
            f1[j] = DS[o];
            r[j] = (C == 1 && J == 1) ? 0.0 : (double)ioh<W>(d)[C * cs + cx(d, i, J, k)];
        }
    }
    e0[threadIdx.x] = f0[0];
    e1[threadIdx.x] = f0[P - 1];
    e2[threadIdx.x] = f1[0];
    e3[threadIdx.x] = f1[P - 1];
    // the component's AMVR / APVR rows in LDS: a wavefront spans several row chunks, so their
    // per-row reads are not wave-uniform
    double *am = lds + 4 * S + 1, *ap = am + (M + 2);
    for (int q = threadIdx.x; q < M + 2; q += S) {
        am[q] = d.amvr[C * (M + 2) + q];
        ap[q] = d.apvr[C * (M + 2) + q];
    }
    __syncthreads();
    const double f0m = sl >= 0 ? e1[sl] : 0.0, f1m = sl >= 0 ? e3[sl] : 1.0;
    const double f0p = sr >= 0 ? e0[sr] : 0.0, f1p = sr >= 0 ? e2[sr] : 1.0;
    double a[P], c[P], ib[P], rho = 0.0;
#pragma unroll
    for (int j = 0; j < P; ++j) {
        const int J = J0 + j;
        double aa = 0.0, cc = 0.0, b = 1.0;
        if (valid && J <= M) {
            const double fm = J > 1 ? (j > 0 ? f0[j - 1] : f0m) : 0.0;
            const double fp = J < M ? (j < P - 1 ? f0[j + 1] : f0p) : 0.0;
            const double dm = J > 1 ? (j > 0 ? f1[j - 1] : f1m) : 1.0;
            const double dq = J < M ? (j < P - 1 ? f1[j + 1] : f1p) : 1.0;
            double amu, apu;
            if constexpr (C != 1) {
                const double coe2 = facoe;
                if (J == 1) {
                    const double muqt = (d.m1[hx(d, i, 1, k)] + d.m1[hx(d, i, 0, k)]) * 0.5;
                    amu = -coe2 * d.amvr1 * muqt;
                    apu = -coe2 * ap[J] * fp;
                    cc = qd<FD>(apu, dq);
                } else if (J == M) {
                    const double muqt = (d.m1[hx(d, i, M, k)] + d.m1[hx(d, i, M + 1, k)]) * 0.5;
                    amu = -coe2 * am[J] * f0[j];
                    apu = -coe2 * d.apvrn * muqt;
                    aa = qd<FD>(amu, dm);
                } else {
                    amu = -coe2 * am[J] * f0[j];
                    apu = -coe2 * ap[J] * fp;
                    aa = qd<FD>(amu, dm);
                    cc = qd<FD>(apu, dq);
                }
            } else {
                const double coe1 = facoe * 4.0 / 3.0;
                amu = -coe1 * am[J] * fm;
                apu = -coe1 * ap[J] * f0[j];
                aa = J == 1 ? 0.0 : qd<FD>(amu, dm);
                cc = qd<FD>(apu, J == 1 ? dq : f1[j]);
            }
            b = 1.0 - qd<FD>(amu + apu, f1[j]);
        }
        a[j] = aa;
        c[j] = cc;
        ib[j] = qd<FD>(1.0, b);
        rho = nanmax(rho, (fabs(aa) + fabs(cc)) * fabs(ib[j]));
    }
    rho = block_max_nonneg(rho, word);
    if (!(rho <= 0.25)) {
        if (valid && tt == 0) io_cn_y_line<C, W>(d, gs, k * L + i, facoe);
        return;
    }
    const int K = jacobi_sweeps(rho);
    double x[P];
#pragma unroll
    for (int j = 0; j < P; ++j) x[j] = r[j] * ib[j];
    for (int it = 0; it < K; ++it) {
        double *w0 = (it & 1) ? e2 : e0, *w1 = (it & 1) ? e3 : e1;   // alternating: one barrier per sweep
        w0[threadIdx.x] = x[0];
        w1[threadIdx.x] = x[P - 1];
        __syncthreads();
        const double xl = sl >= 0 ? w1[sl] : 0.0, xr = sr >= 0 ? w0[sr] : 0.0;
        double xn[P];
#pragma unroll
        for (int j = 0; j < P; ++j) {
            const double xm = j > 0 ? x[j - 1] : xl, xp = j < P - 1 ? x[j + 1] : xr;
            xn[j] = (r[j] - a[j] * xm - c[j] * xp) * ib[j];
        }
#pragma unroll
        for (int j = 0; j < P; ++j) x[j] = xn[j];
    }
    if (valid) {
        double *G = d.g + C * hs;
        const double *G0 = gs + C * hs;
#pragma unroll
        for (int j = 0; j < P; ++j) {
            const int J = J0 + j;
            if (J <= M) {
                const size_t o = hx(d, i, J, k);
                G[o] = (C == 1 && J == 1) ? 0.0 : x[j] + G0[o];
            }
        }
        if (tt == 0) G[hx(d, i, 0, k)] = 0.0;
        if (tt == NT - 1) G[hx(d, i, M + 1, k)] = 0.0;
    }
}

template <typename W, bool FD, int P>
__global__ __launch_bounds__(1024) void k_io_cn_y_jac(IoDev d, const double *__restrict__ gs, double facoe)
{
    extern __shared__ double lds[];
    if (blockIdx.y == 0) io_cn_y_jac<0, P, W, FD>(d, gs, facoe, lds);
    else if (blockIdx.y == 1) io_cn_y_jac<1, P, W, FD>(d, gs, facoe, lds);
    else io_cn_y_jac<2, P, W, FD>(d, gs, facoe, lds);
}

// DIVG_io (DIVG.f90:54-137), iWeightedPre = 0: rows 1..M.
__global__ __launch_bounds__(256) void k_io_divg(IoDev d, double coe0)
{
#pragma clang fp contract(off)
    IO_THREAD(1)
    if (i >= L || J > M) return;
    const int ip = i == L - 1 ? 0 : i + 1, kp = k == N - 1 ? 0 : k + 1;
    const size_t hs = hsz(d), c0 = hx(d, i, J, k), ci = cx(d, i, J, k);
    const double divx = (d.g[hx(d, ip, J, k)] - d.g[c0]) * d.dxi;
    const double divy = (d.g[hs + hx(d, i, J + 1, k)] - d.g[hs + c0]) * d.dyfi[J];
    const double divz = (d.g[2 * hs + hx(d, i, J, kp)] - d.g[2 * hs + c0]) * d.dzi;
    const double dq = divx + divy + divz + d.drho[ci];
    d.rhsp[ci] = dq * coe0;
}

// PRCALC_io (PRESSURE_CORRECTION.f90:102-170; cn: the Crank-Nicolson branch's viscous term
// -BE * RHSLLPHI_io, BE = alpha dt / (2 Re), the Poisson right-hand side) with BC_WALL_PR_io,
// MASSFLUX_UPDATE_io (VELO_CORRECTION.f90:103-179) with BC_WALL_G_io, and VELOCITY_CALC_io:
// rows 0..M+1.
__global__ __launch_bounds__(256) void k_io_correct(IoDev d, double tadt, int cn, double be)
{
#pragma clang fp contract(off)
    IO_THREAD(0)
    if (i >= L || J > M + 1) return;
    const size_t hs = hsz(d), c0 = hx(d, i, J, k);
    if (J >= 1 && J <= M) {   // the wall cells' threads also write the ghost-row copies
        double p = d.pr[c0] + d.dph[c0];
        if (cn) p = p - be * d.rhsp[cx(d, i, J, k)];
        d.pr[c0] = p;
        if (J == 1) d.pr[hx(d, i, 0, k)] = p;
        if (J == M) d.pr[hx(d, i, M + 1, k)] = p;
    }
    if (J == 0 || J == M + 1) {
        for (int c = 0; c < 3; ++c) {
            d.g[c * hs + c0] = 0.0;
            d.q[c * hs + c0] = 0.0;
        }
        return;
    }
    const int im = i == 0 ? L - 1 : i - 1, km = k == 0 ? N - 1 : k - 1;
    const double ph = d.dph[c0];
    const double gx = d.g[c0] - (ph - d.dph[hx(d, im, J, k)]) * (tadt * d.dxi);
    const double gz = d.g[2 * hs + c0] - (ph - d.dph[hx(d, i, J, km)]) * (tadt * d.dzi);
    d.g[c0] = gx;
    d.g[2 * hs + c0] = gz;
    d.q[c0] = gx / d.dstg[c0];
    d.q[2 * hs + c0] = gz / d.dstg[2 * hs + c0];
    if (J == 1) {
        d.g[hs + c0] = 0.0;
        d.q[hs + c0] = 0.0;
    } else {
        const double gy = d.g[hs + c0] - (ph - d.dph[hx(d, i, J - 1, k)]) * (tadt * d.dyci[J]);
        d.g[hs + c0] = gy;
        d.q[hs + c0] = gy / d.dstg[hs + c0];
    }
}

// CFL_io (CFL.f90:59-107): per cell the larger of the convective rate
// (|u_i + u_i+1| DXI + |v_J + v_J+1| DYFI + |w_k + w_k+1| DZI) / 2 and the viscous rate
// 2 CVISC (DXQI + DZQI + DYFI^2) Viscousity / DENSITY; block maxima to d.partial.
__global__ __launch_bounds__(256) void k_io_cfl(IoDev d, const double *__restrict__ dens,
                                                const double *__restrict__ visc)
{
#pragma clang fp contract(off)
    __shared__ double sh[256];
    IO_THREAD(1)
    double v = 0.0;
    if (i < L && J <= M) {
        const int ip = i == L - 1 ? 0 : i + 1, kp = k == N - 1 ? 0 : k + 1;
        const size_t hs = hsz(d), c0 = hx(d, i, J, k);
        const double *Q = d.q;
        const double qcf = (fabs((Q[c0] + Q[hx(d, ip, J, k)]) * d.dxi) +
                            fabs((Q[hs + c0] + Q[hs + hx(d, i, J + 1, k)]) * d.dyfi[J]) +
                            fabs((Q[2 * hs + c0] + Q[2 * hs + hx(d, i, J, kp)]) * d.dzi)) * 0.5;
        const double vcf = 2.0 * d.cvisc * (d.dxqi + d.dzqi + d.dyfi[J] * d.dyfi[J]) * visc[c0] / dens[c0];
        v = nanmax(qcf, vcf);
    }
    const int t = threadIdx.x + 64 * threadIdx.y;
    sh[t] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) sh[t] = nanmax(sh[t], sh[t + o]);
        __syncthreads();
    }
    if (t == 0) d.partial[((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = sh[0];
}

// DIVGCK_io / DIVGCK_Comm_io (DIVGCK.f90:60-175), periodic x: per cell
// |dG1/dx + dG2/dy + dG3/dz + DrhoDtP|; block maxima to d.partial.
__global__ __launch_bounds__(256) void k_io_divck(IoDev d)
{
#pragma clang fp contract(off)
    __shared__ double sh[256];
    IO_THREAD(1)
    double v = 0.0;
    if (i < L && J <= M) {
        const int ip = i == L - 1 ? 0 : i + 1, kp = k == N - 1 ? 0 : k + 1;
        const size_t hs = hsz(d), c0 = hx(d, i, J, k);
        const double divx = (d.g[hx(d, ip, J, k)] - d.g[c0]) * d.dxi;
        const double divy = (d.g[hs + hx(d, i, J + 1, k)] - d.g[hs + c0]) * d.dyfi[J];
        const double divz = (d.g[2 * hs + hx(d, i, J, kp)] - d.g[2 * hs + c0]) * d.dzi;
        v = fabs(divx + divy + divz + d.drho[cx(d, i, J, k)]);
    }
    const int t = threadIdx.x + 64 * threadIdx.y;
    sh[t] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o) sh[t] = nanmax(sh[t], sh[t + o]);
        __syncthreads();
    }
    if (t == 0) d.partial[((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(256) void k_io_reduce_max(const double *__restrict__ part, int n, double *__restrict__ out)
{
    __shared__ double sh[256];
    double v = 0.0;
    for (int j = threadIdx.x; j < n; j += 256) v = nanmax(v, part[j]);
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] = nanmax(sh[threadIdx.x], sh[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

template <typename A, typename B>
__global__ __launch_bounds__(256) void k_io_cvt(const A *__restrict__ a, B *__restrict__ b, size_t n)
{
    for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (size_t)gridDim.x * 256) b[t] = (B)a[t];
}

template <typename A, typename B> hipError_t io_convert(const A *a, B *b, size_t n, hipStream_t st)
{
    hipLaunchKernelGGL((k_io_cvt<A, B>), dim3((unsigned)std::min<size_t>((n + 255) / 256, 65536)), dim3(256), 0, st,
                       a, b, n);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_io_fill(double *__restrict__ p, size_t n, double v)
{
    const size_t j = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (j < n) p[j] = v;
}

#undef IO_THREAD

}  // namespace

struct chs_io {
    int L = 0, M = 0, N = 0;
    chs_ctx *pois = nullptr;      // Poisson solver on the same mesh (its stream is ours)
    hipStream_t stream = nullptr;
    IoDev d{};
    double *buf = nullptr;
    int npart = 0, flow_driven = 1;
    int mom_rows = 4;   // rows per momentum block (4, 2 or 1; the rest of its 4 rows of threads are z-planes)
    int mom_kz = 32;    // z-planes per block of the LDS-staged momentum kernel (0: the per-point kernel)
    int mom_by = 4;     // rows per block of the LDS-staged momentum kernel (4 or 8)
    int mom_nxcd = 8;   // XCDs the LDS-staged momentum kernel groups its tiles for (1: launch order)
    int props_kc = 16;  // z-planes per block of the LDS-staged staggered properties (0: per-point kernel)
    int cnz_cols = 0, cny_cols = 0;   // cn_cols for the z or the y solve alone (0: cn_cols)
    int cnz_p = 8, cny_p = 8;         // line points per thread of the z and y solves (4 or 8; 8:
                                      // 6.21 -> 6.09 (z) and 6.14 (y) ms per heated CN step)
    int cnx_wave = 1;     // x-lines of the Crank-Nicolson solve one per wavefront (L = 64, 128, 256, 512)
    int cn_fastdiv = 1;   // z and y Jacobi coefficients by Newton reciprocals (the sweeps resolve
                          // no finer; 0.26 -> 0.24 and 0.29 -> 0.28 ms per stage at 256 x 128 x 256)
    int cn_cols = 16;   // most x-columns per block of the z and y Jacobi line solves (16: 1.7 %
                        // faster per heated step than 64 at 256 x 128 x 256)
    double *own_d0 = nullptr, *own_m0 = nullptr;   // d.d0 / d.m0 may alias an energy context's
    double *own_m1 = nullptr;                       // d.m1 likewise (Crank-Nicolson only)
    double *cnbuf = nullptr;                        // m1, work, amvr, apvr
    float *wbuf = nullptr;                          // d.explt0f, d.hf (mixed form)
    int pois_fishpack = 0;                          // pressure solve with FISHPACK's (0,0)-mode gauge
    bool dph_valid = false;                         // a substep has solved for DPH
    int mixed = 0;                                  // EXPLT0 and the increment in FP32
    int cn = 0;                                     // iVisScheme
    bool q_fresh = false;   // Q = G / D_STG of the current G and D_STG (written by chs_io_props)
    double cf_given = 0.0, volm = 0.0;
    double tgam[4], troh[4], talp[4];
    std::string err;
};

namespace {

int io_fail(chs_io *m, int code, const std::string &msg)
{
    if (m) m->err = msg;
    g_io_err = msg;
    return code;
}

dim3 io_grid(const chs_io *m, int rows) { return dim3((m->L + 63) / 64, (rows + 3) / 4, m->N); }

int io_hip(chs_io *m, hipError_t e, const char *what)
{
    if (e == hipSuccess) e = hipGetLastError();
    return e == hipSuccess ? CHS_OK : io_fail(m, CHS_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

extern "C" {

int chs_io_create(const chs_coeffs *c, const chs_io_params *p, chs_io **out)
{
    if (!c || !p || !out || c->L < 2 || c->M < 3 || c->N < 2 || (p->flow_driven != 1 && p->flow_driven != 2) ||
        p->gravity < 0 || p->gravity > 3 || !(p->cvisc > 0) || (p->vis_scheme != 0 && p->vis_scheme != 1))
        return io_fail(nullptr, CHS_E_ARG,
                       "chs_io_create: bad argument (flow_driven 1 or 2, gravity 0..3, cvisc > 0, vis_scheme 0 or 1)");
    if (p->vis_scheme == 1 && (c->L < 3 || c->N < 3 || !c->amvr || !c->apvr))
        return io_fail(nullptr, CHS_E_ARG, "chs_io_create: the Crank-Nicolson scheme needs L, N >= 3 and AMVR/APVR");
    *out = nullptr;
    chs_io *m = new chs_io();
    const int L = c->L, M = c->M, N = c->N;
    m->L = L; m->M = M; m->N = N;
    int rc = chs_create(c, CHS_PLAIN_FP64, &m->pois);
    if (rc) {
        std::string e = std::string("chs_io_create: Poisson context: ") + chs_last_error(nullptr);
        delete m;
        return io_fail(nullptr, rc, e);
    }
    m->stream = chs_internal_stream(m->pois);
    const size_t hs = (size_t)N * (M + 2) * L, cs = (size_t)N * M * L;
    m->npart = (int)(((L + 63) / 64) * ((M + 3) / 4) * N);   // partials of the 64 x 4 row-block kernels
    // room for the largest block count of the row-block kernels and of every momentum shape
    const size_t npart_alloc = (size_t)((L + 63) / 64) *
        std::max({(size_t)((M + 3) / 4) * N, (size_t)((M + 1) / 2) * ((N + 1) / 2), (size_t)M * ((N + 3) / 4)});
    const size_t ng = 4 * (size_t)(M + 2);
    const size_t nb = 18 * hs + 9 * cs + ng + npart_alloc + 4;
    hipError_t he = hipMalloc(&m->buf, sizeof(double) * nb);
    if (he == hipSuccess) he = hipMemsetAsync(m->buf, 0, sizeof(double) * nb, m->stream);
    IoDev &d = m->d;
    double *f = m->buf;
    d.g = f; d.g0 = f + 3 * hs; d.q = f + 6 * hs; d.dstg = f + 9 * hs; d.mstg = f + 12 * hs;
    d.pr = f + 15 * hs; d.d0 = f + 16 * hs; d.m0 = f + 17 * hs;
    m->own_d0 = d.d0; m->own_m0 = d.m0;
    f += 18 * hs;
    d.explt0 = f; d.h = f + 3 * cs; d.drho = f + 6 * cs; d.divu = f + 7 * cs; d.rhsp = f + 8 * cs;
    f += 9 * cs;
    d.dph = chs_internal_phi(m->pois);
    double *geo = f;
    d.partial = f + ng;
    d.scal = d.partial + npart_alloc;
    std::vector<double> hg(ng, 0.0);
    for (int J = 1; J <= M; ++J) hg[J] = c->dyfi[J - 1];
    for (int J = 1; J <= M + 1; ++J) {
        hg[(M + 2) + J] = c->dyci[J - 1];
        hg[2 * (M + 2) + J] = c->wff[J - 1];
        hg[3 * (M + 2) + J] = c->wfb[J - 1];
    }
    if (he == hipSuccess) he = hipMemcpyAsync(geo, hg.data(), sizeof(double) * ng, hipMemcpyHostToDevice, m->stream);
    // D_STG = MU_STG = 1 until the first chs_io_props (MEMALLOCT.f90:200-201)
    if (he == hipSuccess) {
        const size_t n = 6 * hs;
        hipLaunchKernelGGL(k_io_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m->stream, d.dstg, n, 1.0);
        he = hipGetLastError();
    }
    if (he == hipSuccess) he = hipStreamSynchronize(m->stream);
    if (he != hipSuccess) {
        chs_io_destroy(m);
        return io_fail(nullptr, CHS_E_HIP, std::string("chs_io_create: ") + hipGetErrorString(he));
    }
    if (p->vis_scheme == 1) {
        // m1 (halo), work (9 cell fields), AMVR / APVR [3][M+2]
        m->cn = 1;
        const size_t na = 3 * (size_t)(M + 2), nc = hs + 9 * cs + 2 * na;
        he = hipMalloc(&m->cnbuf, sizeof(double) * nc);
        std::vector<double> ha(2 * na, 0.0);
        for (int r = 0; r < 3; ++r)
            for (int J = 1; J <= M + 1; ++J) {
                ha[r * (M + 2) + J] = c->amvr[r * (M + 1) + J - 1];
                ha[na + r * (M + 2) + J] = c->apvr[r * (M + 1) + J - 1];
            }
        if (he == hipSuccess) he = hipMemcpyAsync(m->cnbuf + hs + 9 * cs, ha.data(), sizeof(double) * 2 * na,
                                                  hipMemcpyHostToDevice, m->stream);
        if (he == hipSuccess) he = hipStreamSynchronize(m->stream);
        if (he != hipSuccess) {
            chs_io_destroy(m);
            return io_fail(nullptr, CHS_E_HIP, std::string("chs_io_create: ") + hipGetErrorString(he));
        }
        d.m1 = m->own_m1 = m->cnbuf;
        d.work = m->cnbuf + hs;
        d.amvr = m->cnbuf + hs + 9 * cs;
        d.apvr = d.amvr + na;
        // the wall coefficients of the u, w y-stencils (TDMA_COEF.f90:76-87)
        double ucaj = 4.0 / (2.0 / c->dyci[1] + 1.0 / c->dyfi[0]);
        d.amvr1 = ucaj * (2.0 * c->dyfi[0]);
        ucaj = 4.0 / (2.0 / c->dyci[M - 1] + 1.0 / c->dyfi[M - 1]);
        d.apvrn = ucaj * (2.0 * c->dyfi[M - 1]);
    }
    d.dxqi = c->dxqi; d.dzqi = c->dzqi;
    d.L = L; d.M = M; d.N = N;
    d.dxi = c->dxi; d.dzi = c->dzi; d.cvisc = p->cvisc; d.f_a = p->f_a; d.gravity = p->gravity;
    d.dyfi = geo; d.dyci = geo + (M + 2); d.wff = geo + 2 * (M + 2); d.wfb = geo + 3 * (M + 2);
    m->flow_driven = p->flow_driven; m->cf_given = p->cf_given; m->volm = c->volm;
    for (int s = 0; s < 4; ++s) { m->tgam[s] = c->tgam[s]; m->troh[s] = c->troh[s]; m->talp[s] = c->talp[s]; }
    *out = m;
    return CHS_OK;
}

void chs_io_destroy(chs_io *m)
{
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    if (m->buf) (void)hipFree(m->buf);
    if (m->cnbuf) (void)hipFree(m->cnbuf);
    if (m->wbuf) (void)hipFree(m->wbuf);
    if (m->pois) chs_destroy(m->pois);
    delete m;
}

const char *chs_io_last_error(const chs_io *m) { return m ? m->err.c_str() : g_io_err.c_str(); }

int chs_io_upload(chs_io *m, const double *g, const double *g0, const double *pr, const double *d0,
                  const double *m0, const double *explt0)
{
    if (!m || !g || !pr || !d0 || !m0) return io_fail(m, CHS_E_ARG, "chs_io_upload: null argument");
    const size_t hb = sizeof(double) * (size_t)m->N * (m->M + 2) * m->L, cb = sizeof(double) * (size_t)m->N * m->M * m->L;
    m->d.d0 = m->own_d0;
    m->d.m0 = m->own_m0;
    m->q_fresh = false;
    const IoDev &d = m->d;
    hipError_t he = hipMemcpyAsync(d.g, g, 3 * hb, hipMemcpyHostToDevice, m->stream);
    if (he == hipSuccess)
        he = g0 ? hipMemcpyAsync(d.g0, g0, 3 * hb, hipMemcpyHostToDevice, m->stream)
                : hipMemcpyAsync(d.g0, d.g, 3 * hb, hipMemcpyDeviceToDevice, m->stream);
    if (he == hipSuccess) he = hipMemcpyAsync(d.pr, pr, hb, hipMemcpyHostToDevice, m->stream);
    if (he == hipSuccess) he = hipMemcpyAsync(d.d0, d0, hb, hipMemcpyHostToDevice, m->stream);
    if (he == hipSuccess) he = hipMemcpyAsync(d.m0, m0, hb, hipMemcpyHostToDevice, m->stream);
    m->d.m1 = m->own_m1;
    if (he == hipSuccess && m->cn) he = hipMemcpyAsync(m->own_m1, d.m0, hb, hipMemcpyDeviceToDevice, m->stream);
    if (he == hipSuccess)
        he = explt0 ? hipMemcpyAsync(d.explt0, explt0, 3 * cb, hipMemcpyHostToDevice, m->stream)
                    : hipMemsetAsync(d.explt0, 0, 3 * cb, m->stream);
    if (he == hipSuccess && m->mixed) he = io_convert(d.explt0, d.explt0f, 3 * cb / sizeof(double), m->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(m->stream);
    return io_hip(m, he, "chs_io_upload");
}

int chs_io_set_poisson(chs_io *m, int fishpack)
{
    if (!m || (fishpack != 0 && fishpack != 1)) return io_fail(m, CHS_E_ARG, "chs_io_set_poisson: bad argument");
    m->pois_fishpack = fishpack;
    return CHS_OK;
}

int chs_io_set_mixed(chs_io *m, int mixed)
{
    if (!m || (mixed != 0 && mixed != 1)) return io_fail(m, CHS_E_ARG, "chs_io_set_mixed: bad argument");
    const size_t n = (size_t)3 * m->N * m->M * m->L;
    hipError_t he = hipSuccess;
    if (mixed && !m->wbuf) {
        he = hipMalloc(&m->wbuf, sizeof(float) * 2 * n);
        if (he != hipSuccess) {
            m->wbuf = nullptr;
            return io_hip(m, he, "chs_io_set_mixed");
        }
        m->d.explt0f = m->wbuf;
        m->d.hf = m->wbuf + n;
    }
    // the RK history carries over in the new precision
    if (mixed && !m->mixed) he = io_convert(m->d.explt0, m->d.explt0f, n, m->stream);
    if (!mixed && m->mixed) he = io_convert(m->d.explt0f, m->d.explt0, n, m->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(m->stream);
    if (he == hipSuccess) m->mixed = mixed;
    return io_hip(m, he, "chs_io_set_mixed");
}

int chs_io_set_properties(chs_io *m, const double *d0_dev, const double *m0_dev)
{
    if (!m || !d0_dev || !m0_dev) return io_fail(m, CHS_E_ARG, "chs_io_set_properties: null argument");
    const size_t hb = sizeof(double) * (size_t)m->N * (m->M + 2) * m->L;
    m->d.d0 = m->own_d0;
    m->d.m0 = m->own_m0;
    m->q_fresh = false;
    hipError_t he = hipMemcpyAsync(m->d.d0, d0_dev, hb, hipMemcpyDeviceToDevice, m->stream);
    if (he == hipSuccess) he = hipMemcpyAsync(m->d.m0, m0_dev, hb, hipMemcpyDeviceToDevice, m->stream);
    m->d.m1 = m->own_m1;
    if (he == hipSuccess && m->cn) he = hipMemcpyAsync(m->own_m1, m0_dev, hb, hipMemcpyDeviceToDevice, m->stream);
    return io_hip(m, he, "chs_io_set_properties");
}

int chs_io_upload_viscosity(chs_io *m, const double *visc)
{
    if (!m || !visc) return io_fail(m, CHS_E_ARG, "chs_io_upload_viscosity: null argument");
    if (!m->cn) return io_fail(m, CHS_E_ARG, "chs_io_upload_viscosity: the context uses the explicit scheme");
    const size_t hb = sizeof(double) * (size_t)m->N * (m->M + 2) * m->L;
    m->d.m1 = m->own_m1;
    hipError_t he = hipMemcpyAsync(m->own_m1, visc, hb, hipMemcpyHostToDevice, m->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(m->stream);
    return io_hip(m, he, "chs_io_upload_viscosity");
}

int chs_io_props(chs_io *m, int ns)
{
    if (!m || ns < 1 || ns > 3) return io_fail(m, CHS_E_ARG, "chs_io_props: bad argument (NS 1..3)");
    const IoDev &d = m->d;
    auto a16 = [](const void *q) { return ((uintptr_t)q & 15) == 0; };
    const dim3 g = io_grid(m, m->M + 1);
    if (m->props_kc > 0 && m->L % 2 == 0 && a16(d.d0) && a16(d.m0) && a16(d.g) && a16(d.g0)) {
        const int KC = m->props_kc;
        hipLaunchKernelGGL(k_io_props_lds, dim3(g.x, g.y, (m->N + KC - 1) / KC), dim3(64, 4), 0, m->stream, d,
                           m->tgam[ns], m->troh[ns], KC, m->mom_nxcd);
    } else {
        constexpr int KC = 8;
        hipLaunchKernelGGL(k_io_props_z, dim3(g.x, g.y, (m->N + KC - 1) / KC), dim3(64, 4), 0, m->stream, d,
                           m->tgam[ns], m->troh[ns], KC);
    }
    m->q_fresh = true;
    return io_hip(m, hipSuccess, "chs_io_props");
}

int chs_io_set_tuning(chs_io *m, const char *key, int value)
{
    if (!m || !key) return io_fail(m, CHS_E_ARG, "chs_io_set_tuning: bad argument");
    const std::string k(key);
    if (k == "mom_rows") {
        if (value != 1 && value != 2 && value != 4) return io_fail(m, CHS_E_ARG, "chs_io_set_tuning: mom_rows must be 1, 2 or 4");
        m->mom_rows = value;
        return CHS_OK;
    }
    if (k == "mom_kz") {
        if (value < 0 || value > 1024) return io_fail(m, CHS_E_ARG, "chs_io_set_tuning: mom_kz must be 0..1024");
        m->mom_kz = value;
        return CHS_OK;
    }
    if (k == "props_kc") {
        if (value < 0 || value > 1024) return io_fail(m, CHS_E_ARG, "chs_io_set_tuning: props_kc must be 0..1024");
        m->props_kc = value;
        return CHS_OK;
    }
    if (k == "mom_nxcd") {
        if (value < 1 || value > 64) return io_fail(m, CHS_E_ARG, "chs_io_set_tuning: mom_nxcd must be 1..64");
        m->mom_nxcd = value;
        return CHS_OK;
    }
    if (k == "mom_by") {
        if (value != 4 && value != 8) return io_fail(m, CHS_E_ARG, "chs_io_set_tuning: mom_by must be 4 or 8");
        m->mom_by = value;
        return CHS_OK;
    }
    if (k == "cnz_p" || k == "cny_p") {
        if (value != 4 && value != 8) return io_fail(m, CHS_E_ARG, "chs_io_set_tuning: " + k + " must be 4 or 8");
        (k == "cnz_p" ? m->cnz_p : m->cny_p) = value;
        return CHS_OK;
    }
    if (k == "cnz_cols" || k == "cny_cols") {
        if (value < 0 || value > 64) return io_fail(m, CHS_E_ARG, "chs_io_set_tuning: " + k + " must be 0..64");
        (k == "cnz_cols" ? m->cnz_cols : m->cny_cols) = value;
        return CHS_OK;
    }
    if (k == "cn_fastdiv") {
        if (value != 0 && value != 1) return io_fail(m, CHS_E_ARG, "chs_io_set_tuning: cn_fastdiv must be 0 or 1");
        m->cn_fastdiv = value;
        return CHS_OK;
    }
    if (k == "cnx_wave") {
        if (value != 0 && value != 1) return io_fail(m, CHS_E_ARG, "chs_io_set_tuning: cnx_wave must be 0 or 1");
        m->cnx_wave = value;
        return CHS_OK;
    }
    if (k == "cn_cols") {
        if (value < 1 || value > 64) return io_fail(m, CHS_E_ARG, "chs_io_set_tuning: cn_cols must be 1..64");
        m->cn_cols = value;
        return CHS_OK;
    }
    return io_fail(m, CHS_E_ARG, "chs_io_set_tuning: unknown key " + k);
}

}  // extern "C"

namespace {

// The momentum terms and solves of a substep in the work precision W (EXPLT0 and the
// increment): double for the plain form, float for the mixed form.
template <typename W>
void io_momentum_launch(chs_io *m, const double *gs, double tg, double tr, double ta, double dt)
{
    const IoDev &d = m->d;
    const dim3 blk(64, 4), gall = io_grid(m, m->M + 2), gin = io_grid(m, m->M);
    if (!m->q_fresh) hipLaunchKernelGGL(k_io_velocity, gall, blk, 0, m->stream, d, gs);
    m->q_fresh = false;
    // the LDS-staged kernel loads column pairs as 16-byte words: every staged field 16-byte aligned
    auto a16 = [](const void *q) { return q == nullptr || ((uintptr_t)q & 15) == 0; };
    const bool aligned = m->L % 2 == 0 && a16(gs) && a16(d.q) && a16(d.mstg) && a16(d.m0) && a16(d.m1) &&
                         a16(d.pr) && a16(d.divu);
    const int TY = m->mom_rows, TZ = 4 / TY, KZ = aligned ? m->mom_kz : 0, BY = m->mom_by, NX = m->mom_nxcd;
    const dim3 gmom = KZ > 0 ? dim3((m->L + 63) / 64, (m->M + BY - 1) / BY, (m->N + KZ - 1) / KZ)
                             : dim3((m->L + 63) / 64, (m->M + TY - 1) / TY, (m->N + TZ - 1) / TZ);
    const int npart = (int)(gmom.x * gmom.y * gmom.z);   // at most the 64 x 4 row-block count
    if (!m->cn) hipLaunchKernelGGL(k_io_divu, gin, blk, 0, m->stream, d);
    if (KZ > 0) {
        const dim3 bz(64, BY);
        if (m->cn && BY == 4) hipLaunchKernelGGL((k_io_mom_z<true, 4, W>), gmom, bz, 0, m->stream, d, gs, tg, tr, ta, dt, KZ, NX);
        else if (m->cn) hipLaunchKernelGGL((k_io_mom_z<true, 8, W>), gmom, bz, 0, m->stream, d, gs, tg, tr, ta, dt, KZ, NX);
        else if (BY == 4) hipLaunchKernelGGL((k_io_mom_z<false, 4, W>), gmom, bz, 0, m->stream, d, gs, tg, tr, ta, dt, KZ, NX);
        else hipLaunchKernelGGL((k_io_mom_z<false, 8, W>), gmom, bz, 0, m->stream, d, gs, tg, tr, ta, dt, KZ, NX);
    } else if (m->cn) {
        hipLaunchKernelGGL((k_io_mom<true, W>), gmom, blk, 0, m->stream, d, gs, tg, tr, ta, dt, TY);
    } else {
        hipLaunchKernelGGL((k_io_mom<false, W>), gmom, blk, 0, m->stream, d, gs, tg, tr, ta, dt, TY);
    }
    const double *dpg_dev = nullptr;
    double dpg = 0.0;
    if (m->flow_driven == 1) {
        hipLaunchKernelGGL(k_io_reduce, dim3(1), dim3(256), 0, m->stream, d.partial, npart, m->volm, d.scal);
        dpg_dev = d.scal;
    } else {
        dpg = -0.5 * m->cf_given * (ta * dt);
    }
    if (m->cn) {
        // MOMFA_io per component (SOLVERRK3_MOM_IO.f90:61-74); the components are independent
        // here (each reads Q, the properties and its own increment), so they run side by side.
        const double facoe = 0.5 * ta * dt * d.cvisc;
        const int L = m->L, M = m->M, N = m->N;
        const int PX = L / 64;   // x-line points per lane of the wavefront-per-line solve
        if (m->cnx_wave && L % 64 == 0 && (PX == 1 || PX == 2 || PX == 4 || PX == 8)) {
            const dim3 gx((M * N + 3) / 4, 3);
            if (PX == 1) hipLaunchKernelGGL((k_io_cn_x_wave<1, W, false>), gx, dim3(256), 0, m->stream, d, dpg_dev, dpg, facoe);
            else if (PX == 2) hipLaunchKernelGGL((k_io_cn_x_wave<2, W, false>), gx, dim3(256), 0, m->stream, d, dpg_dev, dpg, facoe);
            else if (PX == 4) hipLaunchKernelGGL((k_io_cn_x_wave<4, W, false>), gx, dim3(256), 0, m->stream, d, dpg_dev, dpg, facoe);
            else hipLaunchKernelGGL((k_io_cn_x_wave<8, W, false>), gx, dim3(256), 0, m->stream, d, dpg_dev, dpg, facoe);
        } else if ((L & (L - 1)) == 0 && L >= 4 && L <= 512) {
            const int T = L >= 256 ? 1 : 256 / L;   // x-lines per block
            hipLaunchKernelGGL(k_io_cn_x_jac<W>, dim3((M * N + T - 1) / T, 3), dim3(T * L),
                               sizeof(double) * (2 * T * L + 1), m->stream, d, dpg_dev, dpg, facoe);
        } else {
            hipLaunchKernelGGL(k_io_cn_x<W>, dim3((M * N + 127) / 128, 3), dim3(128), 0, m->stream, d, dpg_dev, dpg,
                               facoe);
        }
        // line points per thread of the z and y kernels (4 or 8)
        const int Pz = m->cnz_p, Py = m->cny_p;
        const int NTz = N / Pz;   // z: block = Tz x-columns x NTz chunks, whole waves
        int Tz = (N % Pz == 0 && NTz >= 2 && NTz <= 1024) ? std::min(m->cnz_cols ? m->cnz_cols : m->cn_cols, 1024 / NTz) : 0;
        if (Tz > 0) Tz -= Tz % (64 / std::gcd(NTz, 64));
        if (Tz > 0) {
            const int ng = (L + Tz - 1) / Tz;
            const dim3 gz(ng * M, 3), bz(Tz * NTz);
            const size_t sz = sizeof(double) * (4 * Tz * NTz + 1);
            if (Pz == 8 && m->cn_fastdiv) hipLaunchKernelGGL((k_io_cn_z_jac<W, true, 8>), gz, bz, sz, m->stream, d, facoe);
            else if (Pz == 8) hipLaunchKernelGGL((k_io_cn_z_jac<W, false, 8>), gz, bz, sz, m->stream, d, facoe);
            else if (m->cn_fastdiv) hipLaunchKernelGGL((k_io_cn_z_jac<W, true, 4>), gz, bz, sz, m->stream, d, facoe);
            else hipLaunchKernelGGL((k_io_cn_z_jac<W, false, 4>), gz, bz, sz, m->stream, d, facoe);
        } else {
            hipLaunchKernelGGL(k_io_cn_z<W>, dim3((M * L + 127) / 128, 3), dim3(128), 0, m->stream, d, facoe);
        }
        const int NTy = (M + Py - 1) / Py;   // y: block = Ty x-columns x NTy chunks, whole waves
        int Ty = NTy <= 1024 ? std::min(m->cny_cols ? m->cny_cols : m->cn_cols, 1024 / NTy) : 0;
        if (Ty > 0) Ty -= Ty % (64 / std::gcd(NTy, 64));
        if (Ty > 0) {
            const int ng = (L + Ty - 1) / Ty;
            const dim3 gy(ng * N, 3), by(Ty * NTy);
            const size_t sy = sizeof(double) * (4 * Ty * NTy + 1 + 2 * (M + 2));
            if (Py == 8 && m->cn_fastdiv) hipLaunchKernelGGL((k_io_cn_y_jac<W, true, 8>), gy, by, sy, m->stream, d, gs, facoe);
            else if (Py == 8) hipLaunchKernelGGL((k_io_cn_y_jac<W, false, 8>), gy, by, sy, m->stream, d, gs, facoe);
            else if (m->cn_fastdiv) hipLaunchKernelGGL((k_io_cn_y_jac<W, true, 4>), gy, by, sy, m->stream, d, gs, facoe);
            else hipLaunchKernelGGL((k_io_cn_y_jac<W, false, 4>), gy, by, sy, m->stream, d, gs, facoe);
        } else {
            hipLaunchKernelGGL(k_io_cn_y<W>, dim3((L * N + 127) / 128, 3), dim3(128), 0, m->stream, d, gs, facoe);
        }
    } else {
        hipLaunchKernelGGL(k_io_update_g<W>, gall, blk, 0, m->stream, d, gs, dpg_dev, dpg);
    }
}

}  // namespace

extern "C" {

int chs_substep_io_momentum(chs_io *m, int ns, double dt)
{
    if (!m || ns < 1 || ns > 3 || !(dt > 0)) return io_fail(m, CHS_E_ARG, "chs_substep_io_momentum: bad argument");
    // G0_io <- G_io (SOLVERRK3_MOM_IO.f90:25) by exchanging the two arrays: the terms read the
    // old G from g0 and k_io_update_g / k_io_cn_y write every row of the new G.
    std::swap(m->d.g, m->d.g0);
    const double *gs = m->d.g0;
    const double tg = m->tgam[ns], tr = m->troh[ns], ta = m->talp[ns];
    if (m->mixed) io_momentum_launch<float>(m, gs, tg, tr, ta, dt);
    else io_momentum_launch<double>(m, gs, tg, tr, ta, dt);
    return io_hip(m, hipSuccess, "chs_substep_io_momentum");
}

int chs_substep_io(chs_io *m, int ns, double dt)
{
    int rc = chs_substep_io_momentum(m, ns, dt);
    if (rc) return rc;
    const IoDev &d = m->d;
    const double ta = m->talp[ns];
    const dim3 blk(64, 4), gall = io_grid(m, m->M + 2), gin = io_grid(m, m->M);
    hipLaunchKernelGGL(k_io_divg, gin, blk, 0, m->stream, d, 1.0 / dt / ta);
    hipError_t he = hipGetLastError();
    if (he != hipSuccess) return io_hip(m, he, "chs_substep_io");
    rc = chs_internal_poisson_in_place(m->pois, d.rhsp, m->pois_fishpack);   // DPH_io is its solution array
    m->dph_valid = true;
    if (rc) return io_fail(m, rc, std::string("chs_substep_io: Poisson: ") + chs_last_error(m->pois));
    hipLaunchKernelGGL(k_io_correct, gall, blk, 0, m->stream, d, dt * ta, m->cn, 0.5 * ta * dt * d.cvisc);
    return io_hip(m, hipSuccess, "chs_substep_io");
}

int chs_io_download(chs_io *m, double *g, double *q, double *pr, double *dph, double *explt0, double *drho,
                    double *rhsp, double *dstg, double *mstg)
{
    if (!m) return io_fail(m, CHS_E_ARG, "chs_io_download: null context");
    const int L = m->L, M = m->M, N = m->N;
    const size_t hs = (size_t)N * (M + 2) * L, cs = (size_t)N * M * L, hb = sizeof(double) * hs, cb = sizeof(double) * cs;
    const IoDev &d = m->d;
    hipError_t he = hipSuccess;
    if (explt0 && m->mixed) he = io_convert(d.explt0f, d.explt0, 3 * cs, m->stream);   // the FP32 history
    struct { double *dst; const double *src; size_t bytes; } cp[] = {
        {g, d.g, 3 * hb}, {q, d.q, 3 * hb}, {pr, d.pr, hb}, {explt0, d.explt0, 3 * cb},
        {drho, d.drho, cb}, {rhsp, d.rhsp, cb}, {dstg, d.dstg, 3 * hb}, {mstg, d.mstg, 3 * hb}};
    for (auto &x : cp)
        if (x.dst && he == hipSuccess) he = hipMemcpyAsync(x.dst, x.src, x.bytes, hipMemcpyDeviceToHost, m->stream);
    if (dph && m->dph_valid && he == hipSuccess) he = hipMemcpyAsync(dph, d.dph, hb, hipMemcpyDeviceToHost, m->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(m->stream);
    if (he != hipSuccess) return io_hip(m, he, "chs_io_download");
    if (dph && !m->dph_valid) std::fill(dph, dph + hs, 0.0);   // no solve yet
    if (dph) {   // ghost rows copy the wall cells (BC_WALL_DPH_io)
        for (int k = 0; k < N; ++k)
            for (int J = 0; J <= M + 1; J += M + 1) {
                const int Js = J == 0 ? 1 : M;
                std::copy(dph + ((size_t)k * (M + 2) + Js) * L, dph + ((size_t)k * (M + 2) + Js + 1) * L,
                          dph + ((size_t)k * (M + 2) + J) * L);
            }
    }
    return CHS_OK;
}

namespace {

int io_cfl(chs_io *m, const double *dens, const double *visc, double dt, double *cfl, const char *what)
{
    const dim3 blk(64, 4), gin = io_grid(m, m->M);
    hipLaunchKernelGGL(k_io_cfl, gin, blk, 0, m->stream, m->d, dens, visc);
    hipLaunchKernelGGL(k_io_reduce_max, dim3(1), dim3(256), 0, m->stream, m->d.partial, m->npart, m->d.scal + 1);
    double v = 0.0;
    hipError_t he = hipGetLastError();
    if (he == hipSuccess) he = hipMemcpyAsync(&v, m->d.scal + 1, sizeof(double), hipMemcpyDeviceToHost, m->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(m->stream);
    if (he != hipSuccess) return io_hip(m, he, what);
    *cfl = v * dt;
    if (!std::isfinite(v)) return io_fail(m, CHS_E_NONFINITE, std::string(what) + ": non-finite velocity or property");
    return CHS_OK;
}

}  // namespace

int chs_max_div_io(chs_io *m, double *maxdiv)
{
    if (!m || !maxdiv) return io_fail(m, CHS_E_ARG, "chs_max_div_io: null argument");
    hipLaunchKernelGGL(k_io_divck, io_grid(m, m->M), dim3(64, 4), 0, m->stream, m->d);
    hipLaunchKernelGGL(k_io_reduce_max, dim3(1), dim3(256), 0, m->stream, m->d.partial, m->npart, m->d.scal + 2);
    double v = 0.0;
    hipError_t he = hipGetLastError();
    if (he == hipSuccess) he = hipMemcpyAsync(&v, m->d.scal + 2, sizeof(double), hipMemcpyDeviceToHost, m->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(m->stream);
    if (he != hipSuccess) return io_hip(m, he, "chs_max_div_io");
    *maxdiv = v;
    if (!std::isfinite(v)) return io_fail(m, CHS_E_NONFINITE, "chs_max_div_io: non-finite divergence");
    return CHS_OK;
}

int chs_cfl_io(chs_io *m, double dt, double *cfl)
{
    if (!m || !cfl) return io_fail(m, CHS_E_ARG, "chs_cfl_io: null argument");
    return io_cfl(m, m->own_d0, m->cn ? m->own_m1 : m->own_m0, dt, cfl, "chs_cfl_io");
}

int chs_heated_cfl(chs_eng *e, chs_io *m, double dt, double *cfl)
{
    if (!e || !m || !cfl) return io_fail(m, CHS_E_ARG, "chs_heated_cfl: null argument");
    hipStream_t es;
    double *ed0, *em0, *em, *eg;
    chs_internal_eng_fields(e, &es, &ed0, &em0, &em, &eg);
    double *ed = chs_internal_eng_density(e);
    hipError_t he = hipStreamSynchronize(es);   // the energy context's properties are final
    if (he != hipSuccess) return io_hip(m, he, "chs_heated_cfl");
    return io_cfl(m, ed, em, dt, cfl, "chs_heated_cfl");
}

int chs_io_sync(chs_io *m)
{
    if (!m) return io_fail(m, CHS_E_ARG, "chs_io_sync: null context");
    return io_hip(m, hipStreamSynchronize(m->stream), "chs_io_sync");
}

int chs_heated_substep(chs_eng *e, chs_io *m, int ns, double dt)
{
    if (!e || !m || ns < 1 || ns > 3 || !(dt > 0)) return io_fail(m, CHS_E_ARG, "chs_heated_substep: bad argument");
    hipStream_t es;
    double *ed0, *em0, *em, *eg;
    chs_internal_eng_fields(e, &es, &ed0, &em0, &em, &eg);
    hipError_t he = hipStreamSynchronize(m->stream);   // G of the last momentum substep is final
    if (he != hipSuccess) return io_hip(m, he, "chs_heated_substep");
    chs_internal_eng_set_g(e, m->d.g);                 // the energy reads the current G in place
    int rc = chs_substep_eng(e, ns, dt);               // synchronous
    if (rc) {
        chs_internal_eng_set_g(e, nullptr);
        return io_fail(m, rc, std::string("chs_heated_substep: energy: ") + chs_eng_last_error(e));
    }
    chs_internal_eng_fields(e, &es, &ed0, &em0, &em, &eg);  // DENSITY0 / Viscousity0 / Viscousity
    m->d.d0 = ed0;                                     // read in place by the momentum kernels
    m->d.m0 = em0;
    if (m->cn) m->d.m1 = em;
    rc = chs_io_props(m, ns);
    if (!rc) rc = chs_substep_io(m, ns, dt);
    // the launches above hold the aliases by value; later calls see each context's own arrays
    chs_internal_eng_set_g(e, nullptr);
    m->d.d0 = m->own_d0;
    m->d.m0 = m->own_m0;
    m->d.m1 = m->own_m1;
    return rc;
}

}  // extern "C"
