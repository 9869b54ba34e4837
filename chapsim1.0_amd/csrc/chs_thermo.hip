// Thermophysical property tables for the heated channel (table mode, iThermoProperty =
// search_table): host restatement of the reference's table set-up and a GPU kernel for the
// per-cell update THERM_PROP_UPDATE_FROM_DH (THERM_PROP_UPDATE.f90:14-67).
//
// Bit-exactness: every lookup is a bisection (integer) followed by a Horner evaluation in
// FP64 with the reference's operation order and no fused multiply-add (contraction is
// switched off below), so results equal the reference build's bit for bit; the table set-up
// is restated in the same operation order for the same reason.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <utility>
#include <vector>

#include "../../include/chapsim_hip.h"

namespace {

std::string g_prop_err;

// ---------------------------------------------------------------------------------------
// Device
// ---------------------------------------------------------------------------------------
struct SplineDev {
    const double *x, *y, *b, *c, *d;
    int imin, imax;   // 1-based, as IMIN_*/IMAX_* of the reference
};

// Bucket index over a non-decreasing searched table x: bucket b = floor((xv - x0) inv),
// clamped to 0..nb-1, holds an index range lohi[2b] .. lohi[2b+1] that contains the
// bisection's result for every xv of the bucket (the result, the count of x <= xv clamped to
// 1..n, is monotone in xv; the range has one index of margin either side for the rounding of
// the bucket edges).  Bisection inside that range gives the same index in one or two steps
// instead of log2(n).  nb = 0: no index (table not monotone), full bisection.
struct BucketDev {
    const int *lohi;
    double x0, inv;
    int nb;
};

struct PropDev {
    int n;
    SplineDev dhh, ht, hd, hm, hk, hcp, hb;
    BucketDev bdh, bh;   // over LIST_DH (dhh.x) and LIST_H (ht.x and the other H splines)
    // The same coefficients as records, one per table index, for the property update's
    // evaluation (a cell's reads of one index then come in a few 16-byte loads):
    // recd[8 i] = LIST_DH(i), then H's y, b, c, d (2 doubles of padding);
    // rech[24 i] = LIST_H(i), then y, b, c, d of the H->T, H->M, H->K, H->D, H->Cp splines
    // (3 of padding).
    const double *recd, *rech;
};

// The bisection of spline_interpolation_* (I = 1, J = N+1, thermo_properties_table_searching.f90:
// 700-712), started from the bucket's index range when there is one.
__device__ __forceinline__ int spl_index_b(const BucketDev &bk, const double *x, int n, double xv)
{
    int I = 1, J = n + 1;
    if (bk.nb > 0) {
        const double f = (xv - bk.x0) * bk.inv;
        if (f == f) {   // a NaN searches the whole table, as the plain bisection
            const int b = f <= 0.0 ? 0 : (f >= (double)(bk.nb - 1) ? bk.nb - 1 : (int)f);
            const int2 r = reinterpret_cast<const int2 *>(bk.lohi)[b];
            I = r.x;
            J = r.y + 1;
        }
    }
    while (J > I + 1) {
        const int K = (I + J) / 2;
        if (xv < x[K - 1]) J = K;
        else I = K;
    }
    return I;
}

// spline_interpolation_* (thermo_properties_table_searching.f90:678-723 and siblings):
// range check (tolerance 1e-8), bisection I = 1, J = N+1, then
// Y(I) + DX*(B(I) + DX*(C(I) + DX*D(I))).
__device__ double spl_eval(const SplineDev &s, const BucketDev &bk, int n, double xv, int *err, int bit)
{
#pragma clang fp contract(off)
    if ((xv - s.x[s.imin - 1]) < -1.0e-8) atomicOr(err, bit);
    if ((xv - s.x[s.imax - 1]) > 1.0e-8) atomicOr(err, bit << 8);
    const int I = spl_index_b(bk, s.x, n, xv);
    const double dx = xv - s.x[I - 1];
    return s.y[I - 1] + dx * (s.b[I - 1] + dx * (s.c[I - 1] + dx * s.d[I - 1]));
}

// THERM_PROP_UPDATE_FROM_DH, table branch (THERM_PROP_UPDATE.f90:45-52), one thread per
// value.  D is taken from the H->M spline, M from H->K and K from H->D, as the reference
// does (lines 49-51).  out = [H | T | D | M | K | Cp | B], each n values.
__global__ __launch_bounds__(256) void k_prop_from_dh(PropDev p, long n, const double *__restrict__ dh,
                                                      double *__restrict__ out, int *__restrict__ err)
{
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    const double h = spl_eval(p.dhh, p.bdh, p.n, dh[t], err, 1);
    out[t] = h;
    out[n + t] = spl_eval(p.ht, p.bh, p.n, h, err, 2);
    out[2 * n + t] = spl_eval(p.hm, p.bh, p.n, h, err, 4);
    out[3 * n + t] = spl_eval(p.hk, p.bh, p.n, h, err, 8);
    out[4 * n + t] = spl_eval(p.hd, p.bh, p.n, h, err, 16);
    out[5 * n + t] = spl_eval(p.hcp, p.bh, p.n, h, err, 32);
    out[6 * n + t] = spl_eval(p.hb, p.bh, p.n, h, err, 64);
}

// ---------------------------------------------------------------------------------------
// Host set-up restatement
// ---------------------------------------------------------------------------------------
// CUBIC_SPLINE (thermo_properties_table_searching.f90:842-936), Forsythe, Malcolm & Moler's
// natural-end spline with third derivatives matched by divided differences.
void cubic_spline(int N, const double *X, const double *Y, double *B, double *C, double *D)
{
#pragma clang fp contract(off)
    if (N < 2) return;
    if (N < 3) {
        B[0] = (Y[1] - Y[0]) / (X[1] - X[0]);
        C[0] = 0.0; D[0] = 0.0;
        B[1] = B[0]; C[1] = 0.0; D[1] = 0.0;
        return;
    }
    // 1-based views
    const double *x = X - 1, *y = Y - 1;
    double *b = B - 1, *c = C - 1, *d = D - 1;
    const int nm1 = N - 1;
    d[1] = x[2] - x[1];
    c[2] = (y[2] - y[1]) / d[1];
    for (int i = 2; i <= nm1; ++i) {
        d[i] = x[i + 1] - x[i];
        b[i] = 2.0 * (d[i - 1] + d[i]);
        c[i + 1] = (y[i + 1] - y[i]) / d[i];
        c[i] = c[i + 1] - c[i];
    }
    b[1] = -d[1];
    b[N] = -d[N - 1];
    c[1] = 0.0;
    c[N] = 0.0;
    if (N != 3) {
        c[1] = c[3] / (x[4] - x[2]) - c[2] / (x[3] - x[1]);
        c[N] = c[N - 1] / (x[N] - x[N - 2]) - c[N - 2] / (x[N - 1] - x[N - 3]);
        c[1] = c[1] * d[1] * d[1] / (x[4] - x[1]);
        c[N] = -c[N] * (d[N - 1] * d[N - 1]) / (x[N] - x[N - 3]);
    }
    for (int i = 2; i <= N; ++i) {
        const double t = d[i - 1] / b[i - 1];
        b[i] = b[i] - t * d[i - 1];
        c[i] = c[i] - t * c[i - 1];
    }
    c[N] = c[N] / b[N];
    for (int l = 1; l <= nm1; ++l) {
        const int i = N - l;
        c[i] = (c[i] - d[i] * c[i + 1]) / b[i];
    }
    b[N] = (y[N] - y[nm1]) / d[nm1] + d[nm1] * (c[nm1] + 2.0 * c[N]);
    for (int i = 1; i <= nm1; ++i) {
        b[i] = (y[i + 1] - y[i]) / d[i] - d[i] * (c[i + 1] + 2.0 * c[i]);
        d[i] = (c[i + 1] - c[i]) / d[i];
        c[i] = 3.0 * c[i];
    }
    c[N] = 3.0 * c[N];
    d[N] = d[nm1];
}

int first_max(const std::vector<double> &a)
{
    int k = 0;
    for (int i = 1; i < (int)a.size(); ++i)
        if (a[i] > a[k]) k = i;
    return k + 1;
}
int first_min(const std::vector<double> &a)
{
    int k = 0;
    for (int i = 1; i < (int)a.size(); ++i)
        if (a[i] < a[k]) k = i;
    return k + 1;
}

// ---------------------------------------------------------------------------------------
// Energy equation of the main domain: SOLVERRK3_ENG_io (SOLVERRK3_ENG_IO.f90:11-82).  State as
// halo fields [N][M+2][L] (the reference's (L, 0:M+1, N)); one y-domain; channel (RCCI1 =
// RNDI1 = RCCI2 = 1, COORDJR.f90:39-42).  Every expression keeps the reference's operation
// order, contraction off, so results follow the reference's rounding.
struct EngDev {
    int L, M, N;
    double dxi, dzi, dxqi, dzqi;
    const double *dyfi, *dyci, *wff, *wfb;   // indexed by the reference's J (dyfi 1..M, others 1..M+1)
    double *dh, *h, *t, *d, *m, *k, *cp;     // halo fields
    const double *g1, *g2, *g3;              // G_io components, halo fields
    double *rhs0;                            // cells [N][M][L]
    const double *wallflux;                  // [2][N][L]
    int wall_type0, wall_type1, top_face;
    double cthecd, dhmin, dhmax;
    double ws[2][5];                         // fixed-temperature wall states T, H, D, M, K
};

__device__ __forceinline__ size_t ehx(const EngDev &e, int i, int J, int k)
{
    return ((size_t)k * (e.M + 2) + J) * e.L + i;
}

// Row-J coefficients of RHS_ENERGY_EXPLICIT (RHS_ENERGY_EXPLICIT.f90:43-150).
struct EngRow {
    double coe1, coe11, coe2, coe3, coe33, coe22, coe221, coe222, wffp, wfbp, wff, wfb;
};
__device__ __forceinline__ EngRow eng_row(const EngDev &e, int J)
{
#pragma clang fp contract(off)
    const int M = e.M;
    const double one = 1.0, xnd2cl = 0.5, znd2cl = 0.5;
    EngRow c;
    c.coe1 = e.dxi * xnd2cl * 0.5;
    c.coe11 = e.dxqi * xnd2cl * e.cthecd;
    c.coe2 = e.dyfi[J] * one * 0.5;
    c.coe3 = e.dzi * znd2cl * one * 0.5;
    c.coe33 = e.dzqi * znd2cl * one * e.cthecd;
    c.coe22 = e.dyfi[J] * one * e.cthecd;
    if (J == 1) {
        c.coe221 = e.dyci[J + 1] / one;
        c.coe222 = e.wall_type0 == 1 ? 1.0 / one : e.dyci[J] / one;
    } else if (J == M) {
        c.coe222 = e.dyci[J] / one;
        c.coe221 = e.wall_type1 == 1 ? 1.0 / one : e.dyci[J + 1] / one;
    } else {
        c.coe221 = e.dyci[J + 1] / one;
        c.coe222 = e.dyci[J] / one;
    }
    c.wffp = e.wff[J + 1]; c.wfbp = e.wfb[J + 1]; c.wff = e.wff[J]; c.wfb = e.wfb[J];
    return c;
}

// Fields the energy stencil reads in plane k: X(f, di, dj) = f at (i + di, J + dj, k).
enum EngF { EH, EK, ET, EG1, EG2 };

// R of RHS_ENERGY_EXPLICIT at (i, J, k) of row J < or = M: convection of H with G and the
// conduction K grad T; the z-neighbours come as values (m: k-1, c: k, p: k+1), wfl is the
// iso-flux wall's heat flux at (i, k) where row J is a wall row of that kind.
template <typename A>
__device__ __forceinline__ double eng_rhs_point(const EngDev &e, const EngRow &c, const A &X, int J, double Hm,
                                                double Hc, double Hp, double Km, double Kc, double Kp, double Tm,
                                                double Tc, double Tp, double G3c, double G3p, double wfl)
{
#pragma clang fp contract(off)
    const int M = e.M;
    double dx_dhu = (X(EH, 1, 0) + Hc) * (X(EG1, 1, 0) + X(EG1, 1, 0)) - (Hc + X(EH, -1, 0)) * (X(EG1, 0, 0) + X(EG1, 0, 0));
    dx_dhu = dx_dhu * c.coe1;
    double dy_dhv = (c.wffp * X(EH, 0, 1) + c.wfbp * Hc) * (X(EG2, 0, 1) + X(EG2, 0, 1)) -
                    (c.wff * Hc + c.wfb * X(EH, 0, -1)) * (X(EG2, 0, 0) + X(EG2, 0, 0));
    dy_dhv = dy_dhv * c.coe2;
    double dz_dhw = (Hp + Hc) * (G3p + G3p) - (Hc + Hm) * (G3c + G3c);
    dz_dhw = dz_dhw * c.coe3;
    double dx_kdt = (X(EK, 1, 0) + Kc) * (X(ET, 1, 0) - Tc) - (Kc + X(EK, -1, 0)) * (Tc - X(ET, -1, 0));
    dx_kdt = dx_kdt * c.coe11;
    double kdty1 = (c.wffp * X(EK, 0, 1) + c.wfbp * Kc) * (X(ET, 0, 1) - Tc) * c.coe221;
    double kdty2 = (c.wff * Kc + c.wfb * X(EK, 0, -1)) * (Tc - X(ET, 0, -1)) * c.coe222;
    if (J == 1 && e.wall_type0 == 1)
        kdty2 = -wfl * c.coe222;
    else if (J == M && e.wall_type1 == 1)
        kdty1 = -wfl * c.coe221;
    const double dy_kdt = (kdty1 - kdty2) * c.coe22;
    double dz_kdt = (Kp + Kc) * (Tp - Tc) - (Km + Kc) * (Tc - Tm);
    dz_kdt = dz_kdt * c.coe33;
    return -dx_dhu - dy_dhv - dz_dhw + dx_kdt + dy_kdt + dz_kdt;
}

// The iso-flux wall's heat flux the point's row reads (0 where the row is no such wall).
__device__ __forceinline__ double eng_wall_flux(const EngDev &e, int i, int J, int k)
{
    if (J == 1 && e.wall_type0 == 1) return e.wallflux[(size_t)k * e.L + i];
    if (J == e.M && e.wall_type1 == 1) return e.wallflux[((size_t)e.N + k) * e.L + i];
    return 0.0;
}

// The RK update with the table limiter (SOLVERRK3_ENG_IO.f90:35-47): DH += ca R + cb R0,
// clamp, R0 = R.
__device__ __forceinline__ void eng_update(const EngDev &e, size_t h0, size_t c0, double r, double dh0, double r0,
                                           double ca, double cb)
{
#pragma clang fp contract(off)
    double dh = dh0 + ca * r + cb * r0;
    dh = fmin(e.dhmax - 1.0e-14, dh);   // REALMIN (modules.f90:56)
    dh = fmax(e.dhmin + 1.0e-14, dh);
    e.dh[h0] = dh;
    e.rhs0[c0] = r;
}

// RHS_ENERGY_EXPLICIT (RHS_ENERGY_EXPLICIT.f90:43-150) and the RK update, one thread per
// (i, J) marching KC z-planes: H, K, T at (i, J) of plane k and the plane-(k+1) values of H,
// K, T, G_z read for plane k's z-differences stay in registers for the next plane.
__global__ __launch_bounds__(256) void k_eng_rhs_update(EngDev e, double ca, double cb, int KC)
{
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int J = blockIdx.y * 4 + threadIdx.y + 1;
    if (i >= e.L || J > e.M) return;
    const int L = e.L, N = e.N, M = e.M;
    const int k0 = blockIdx.z * KC, k1 = min(N, k0 + KC);
    const int ip = (i == L - 1) ? 0 : i + 1, im = (i == 0) ? L - 1 : i - 1;
    const EngRow c = eng_row(e, J);
    const double *H = e.h, *T = e.t, *K = e.k;
    auto F = [&](const double *f, int ii, int jj, int kk) { return f[ehx(e, ii, jj, kk)]; };
    const int kb = k0 == 0 ? N - 1 : k0 - 1;
    double Hm = F(H, i, J, kb), Km = F(K, i, J, kb), Tm = F(T, i, J, kb);
    double Hc = F(H, i, J, k0), Kc = F(K, i, J, k0), Tc = F(T, i, J, k0), G3c = F(e.g3, i, J, k0);
    for (int k = k0; k < k1; ++k) {
        const int kp = (k == N - 1) ? 0 : k + 1;
        const double Hp = F(H, i, J, kp), Kp = F(K, i, J, kp), Tp = F(T, i, J, kp), G3p = F(e.g3, i, J, kp);
        auto X = [&](int f, int di, int dj) {
            const double *a = f == EH ? H : f == EK ? K : f == ET ? T : f == EG1 ? e.g1 : e.g2;
            return F(a, di > 0 ? ip : (di < 0 ? im : i), J + dj, k);
        };
        const double r = eng_rhs_point(e, c, X, J, Hm, Hc, Hp, Km, Kc, Kp, Tm, Tc, Tp, G3c, G3p, eng_wall_flux(e, i, J, k));
        const size_t c0 = ((size_t)k * M + (J - 1)) * L + i;
        const size_t h0 = ehx(e, i, J, k);
        eng_update(e, h0, c0, r, e.dh[h0], e.rhs0[c0], ca, cb);
        Hm = Hc; Km = Kc; Tm = Tc;
        Hc = Hp; Kc = Kp; Tc = Tp; G3c = G3p;
    }
}

// The same from LDS: a 64 x BY block owns columns i0..i0+63 and rows J0..J0+BY-1 and marches
// KC z-planes; H, K, T, G_x, G_y are staged per plane as (BY + 2) x 68 tiles (columns
// i0-2 .. i0+65, x periodic, loaded as 16-byte column pairs) in a ring of two plane slots,
// plane k-1's H, K, T stay in registers, and the point's DH, R0, G_z and wall flux are loaded a
// plane ahead of the tile loads.  Identical arithmetic to k_eng_rhs_update.
__device__ __forceinline__ int eng_xcd_tile(int p, int nb, int nxcd)
{
    if (nxcd <= 1) return p;
    const int xcd = p % nxcd, slot = p / nxcd, per = nb / nxcd, rem = nb % nxcd;
    return xcd < rem ? xcd * (per + 1) + slot : rem * (per + 1) + (xcd - rem) * per + slot;
}

template <int BY>
__global__ __launch_bounds__(64 * BY) void k_eng_rhs_lds(EngDev e, double ca, double cb, int KC, int nxcd)
{
    constexpr int NT = 64 * BY, TW = 68, TS = (BY + 2) * TW, NF = 5, NCH = TS / 2;
    static_assert(NCH <= NT, "one column pair per thread");
    __shared__ __align__(16) double tiles[2][NF * TS];
    const int L = e.L, M = e.M, N = e.N;
    const int tx = threadIdx.x, ty = __builtin_amdgcn_readfirstlane(threadIdx.y), t = ty * 64 + tx;
    const int gx = gridDim.x, gy = gridDim.y;
    const int tile = eng_xcd_tile(blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z), gx * gy * gridDim.z, nxcd);
    const int bx = tile % gx, by = (tile / gx) % gy, bz = tile / (gx * gy);
    const int i0 = bx * 64, J0 = 1 + by * BY, k0 = bz * KC, k1 = min(N, k0 + KC);
    const int i = i0 + tx, J = J0 + ty;
    const bool act = i < L && J <= M;
    const size_t ph = (size_t)(M + 2) * L;
    // the column pair this thread stages (rows past M + 1 read row M + 1 instead)
    const bool stg = t < NCH;
    unsigned oh;
    {
        const int n = stg ? t : 0, r = n / (TW / 2), q = n % (TW / 2), jj = J0 - 1 + r;
        oh = (unsigned)(min(jj, M + 1) * L + ((i0 - 2 + 2 * q) % L + L) % L);
    }
    const double *base[NF] = {e.h, e.k, e.t, e.g1, e.g2};
    double pre0[NF], pre1[NF];
    auto load = [&](int kk) {
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            const double2 v = *reinterpret_cast<const double2 *>(base[f] + (size_t)kk * ph + oh);
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
    const int Jr = min(J, M), ir = act ? i : 0;
    const EngRow c = eng_row(e, Jr);
    // the point's own values a plane ahead: DH, R0, G_z of the next plane, the wall flux
    double own[4] = {0.0, 0.0, 0.0, 0.0}, ownn[4] = {0.0, 0.0, 0.0, 0.0};
    auto loadown = [&](int kk, double (&o)[4]) {
        if (act) {
            o[0] = e.dh[ehx(e, i, J, kk)];
            o[1] = e.rhs0[((size_t)kk * M + (J - 1)) * L + i];
            o[2] = e.g3[ehx(e, i, J, kk == N - 1 ? 0 : kk + 1)];
            o[3] = eng_wall_flux(e, i, J, kk);
        }
    };
    const int kb = k0 == 0 ? N - 1 : k0 - 1;
    double Hm = e.h[ehx(e, ir, Jr, kb)], Km = e.k[ehx(e, ir, Jr, kb)], Tm = e.t[ehx(e, ir, Jr, kb)];
    double G3c = e.g3[ehx(e, ir, Jr, k0)];
    loadown(k0, own);
    load(k0);
    store(0);
    load(k0 + 1 == N ? 0 : k0 + 1);
    store(1);
    __syncthreads();
    const int cell = (ty + 1) * TW + tx + 2;
    for (int k = k0; k < k1; ++k) {
        const int cur = (k - k0) & 1;
        const bool more = k + 1 < k1;
        if (more) {
            loadown(k + 1, ownn);
            load(k + 2 >= N ? k + 2 - N : k + 2);
        }
        const double *c0p = &tiles[cur][cell], *c1p = &tiles[cur ^ 1][cell];
        const double Hc = c0p[EH * TS], Kc = c0p[EK * TS], Tc = c0p[ET * TS];
        const double Hp = c1p[EH * TS], Kp = c1p[EK * TS], Tp = c1p[ET * TS];
        if (act) {
            auto X = [&](int f, int di, int dj) { return c0p[f * TS + dj * TW + di]; };
            const double r = eng_rhs_point(e, c, X, J, Hm, Hc, Hp, Km, Kc, Kp, Tm, Tc, Tp, G3c, own[2], own[3]);
            eng_update(e, ehx(e, i, J, k), ((size_t)k * M + (J - 1)) * L + i, r, own[0], own[1], ca, cb);
        }
        Hm = Hc; Km = Kc; Tm = Tc; G3c = own[2];
#pragma unroll
        for (int q = 0; q < 4; ++q) own[q] = ownn[q];
        __syncthreads();
        if (more) store(cur);
        __syncthreads();
    }
}

__device__ __forceinline__ double spl_at(const SplineDev &s, int I, double xv)
{
#pragma clang fp contract(off)
    const double dx = xv - s.x[I - 1];
    return s.y[I - 1] + dx * (s.b[I - 1] + dx * (s.c[I - 1] + dx * s.d[I - 1]));
}

// spl_at from a record's x and y, b, c, d (the same expression).
__device__ __forceinline__ double spl_rec(double x, double y, double b, double c, double d, double xv)
{
#pragma clang fp contract(off)
    const double dx = xv - x;
    return y + dx * (b + dx * (c + dx * d));
}

// spl_index_b for C values in lockstep: the bisection steps are branch-free (a finished search
// re-reads its own entry and keeps its bounds), so the C searches' loads go out together.
// Two lockstep steps cover a bucket's index range unless the table clusters; the rest, if
// any, runs as the plain loop.  Each result equals spl_index_b's.
template <int C>
__device__ __forceinline__ void spl_index_lock(const BucketDev &bk, const double *x, int n, const double *xv,
                                               int *I)
{
    int J[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        I[c] = 1;
        J[c] = n + 1;
        if (bk.nb > 0) {
            const double f = (xv[c] - bk.x0) * bk.inv;
            if (f == f) {
                const int b = f <= 0.0 ? 0 : (f >= (double)(bk.nb - 1) ? bk.nb - 1 : (int)f);
                const int2 r = reinterpret_cast<const int2 *>(bk.lohi)[b];
                I[c] = r.x;
                J[c] = r.y + 1;
            }
        }
    }
#pragma unroll
    for (int st = 0; st < 2; ++st) {
        double xk[C];
        int K[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            K[c] = (I[c] + J[c]) / 2;   // = I when the search is done (J = I + 1): a valid entry
            xk[c] = x[K[c] - 1];
        }
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const bool go = J[c] > I[c] + 1, lt = xv[c] < xk[c];
            J[c] = (go && lt) ? K[c] : J[c];
            I[c] = (go && !lt) ? K[c] : I[c];
        }
    }
#pragma unroll
    for (int c = 0; c < C; ++c)
        while (J[c] > I[c] + 1) {
            const int K = (I[c] + J[c]) / 2;
            if (xv[c] < x[K - 1]) J[c] = K;
            else I[c] = K;
        }
}

// THERM_PROP_UPDATE_FROM_DH for C cells (h0[c], those with act[c]).  The five H-based splines
// share their table (LIST_H) and range (IMIN_H, IMAX_H), so one search serves all of them;
// each result equals spl_eval's.  xdh / xh: LIST_DH / LIST_H (LDS copies or the global
// tables).  An inactive cell searches the table's first entry and stores nothing.
template <int C>
__device__ void eng_cells_props(const EngDev &e, const PropDev &p, const size_t *h0, const bool *act, int *err,
                                const double *xdh, const double *xh)
{
    double dv[C], hv[C];
    int I[C];
#pragma unroll
    for (int c = 0; c < C; ++c) dv[c] = act[c] ? e.dh[h0[c]] : p.dhh.x[0];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        if (!act[c]) continue;
        if ((dv[c] - p.dhh.x[p.dhh.imin - 1]) < -1.0e-8) atomicOr(err, 1);
        if ((dv[c] - p.dhh.x[p.dhh.imax - 1]) > 1.0e-8) atomicOr(err, 1 << 8);
    }
    spl_index_lock<C>(p.bdh, xdh, p.n, dv, I);
#pragma unroll
    for (int c = 0; c < C; ++c) {
        hv[c] = 0.0;   // an inactive cell searches from any value
        if (act[c]) {
            const double2 *r = reinterpret_cast<const double2 *>(p.recd + (size_t)8 * (I[c] - 1));
            const double2 r0 = r[0], r1 = r[1], r2 = r[2];
            hv[c] = spl_rec(r0.x, r0.y, r1.x, r1.y, r2.x, dv[c]);
        }
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
        if (!act[c]) continue;
        if ((hv[c] - p.ht.x[p.ht.imin - 1]) < -1.0e-8) atomicOr(err, 2 | 4 | 8 | 16 | 32);
        if ((hv[c] - p.ht.x[p.ht.imax - 1]) > 1.0e-8) atomicOr(err, (2 | 4 | 8 | 16 | 32) << 8);
    }
    spl_index_lock<C>(p.bh, xh, p.n, hv, I);
#pragma unroll
    for (int c = 0; c < C; ++c) {
        if (!act[c]) continue;
        const size_t o = h0[c];
        const double2 *r = reinterpret_cast<const double2 *>(p.rech + (size_t)24 * (I[c] - 1));
        double2 v[11];
#pragma unroll
        for (int q = 0; q < 11; ++q) v[q] = r[q];
        const double x = v[0].x, h = hv[c];
        e.h[o] = h;
        e.t[o] = spl_rec(x, v[0].y, v[1].x, v[1].y, v[2].x, h);
        e.d[o] = spl_rec(x, v[2].y, v[3].x, v[3].y, v[4].x, h);    // D from H->M, as the reference
        e.m[o] = spl_rec(x, v[4].y, v[5].x, v[5].y, v[6].x, h);    // M from H->K
        e.k[o] = spl_rec(x, v[6].y, v[7].x, v[7].y, v[8].x, h);    // K from H->D
        e.cp[o] = spl_rec(x, v[8].y, v[9].x, v[9].y, v[10].x, h);
    }
}

__device__ void eng_cell_props(const EngDev &e, const PropDev &p, size_t h0, int *err, const double *xdh,
                               const double *xh)
{
    const bool act = true;
    eng_cells_props<1>(e, p, &h0, &act, err, xdh, xh);
}

// THERM_PROP_UPDATE (THERM_PROP_UPDATE.f90:124-168) over rows j0..j1.  With lds != 0 the two
// searched tables (2 n doubles of dynamic LDS) are staged per block first.
__global__ __launch_bounds__(256) void k_eng_props(EngDev e, PropDev p, int j0, int j1, int *err, int lds)
{
    extern __shared__ double tab[];
    const double *xdh = p.dhh.x, *xh = p.ht.x;
    if (lds) {
        for (int t = threadIdx.x + 64 * threadIdx.y; t < p.n; t += 256) {
            tab[t] = p.dhh.x[t];
            tab[p.n + t] = p.ht.x[t];
        }
        __syncthreads();
        xdh = tab;
        xh = tab + p.n;
    }
    // two cells per thread, z-planes k and k + gridDim.z, searched in lockstep
    const int i = blockIdx.x * 64 + threadIdx.x;
    const int J = j0 + blockIdx.y * 4 + threadIdx.y;
    const int k = blockIdx.z, k2 = k + gridDim.z;
    if (i >= e.L || J > j1) return;
    const size_t h0[2] = {ehx(e, i, J, k), ehx(e, i, J, k2 < e.N ? k2 : k)};
    const bool act[2] = {true, k2 < e.N};
    eng_cells_props<2>(e, p, h0, act, err, xdh, xh);
}

// BC_WALL_THERMAL(IMAIND) (BC_WALL.f90:622-643): iso-flux walls extrapolate DH to the ghost
// row and look up H, T, D, M, K, Cp there (BC_WALL_Isoflux :575-615); iso-thermal walls take
// the wall state and DH = H D (BC_WALL_IsoTHERMAL :484-538).  One thread per (i, k).
__global__ __launch_bounds__(256) void k_eng_walls(EngDev e, PropDev p, int *err)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int k = blockIdx.y;
    if (i >= e.L) return;
    const int M = e.M;
    for (int w = 0; w < 2; ++w) {
        const int J = w == 0 ? 0 : M + 1;
        const size_t h0 = ehx(e, i, J, k);
        if ((w == 0 ? e.wall_type0 : e.wall_type1) == 1) {
            if (w == 0)
                e.dh[h0] = 2.0 * e.dh[ehx(e, i, 1, k)] - e.wff[2] * e.dh[ehx(e, i, 2, k)] -
                           e.wfb[2] * e.dh[ehx(e, i, 1, k)];
            else
                e.dh[h0] = 2.0 * e.dh[ehx(e, i, M, k)] - e.wff[e.top_face] * e.dh[ehx(e, i, M, k)] -
                           e.wfb[e.top_face] * e.dh[ehx(e, i, M - 1, k)];
            eng_cell_props(e, p, h0, err, p.dhh.x, p.ht.x);
        } else {
            e.t[h0] = e.ws[w][0];
            e.h[h0] = e.ws[w][1];
            e.d[h0] = e.ws[w][2];
            e.m[h0] = e.ws[w][3];
            e.k[h0] = e.ws[w][4];
            e.dh[h0] = e.h[h0] * e.d[h0];
        }
    }
}

}  // namespace

struct chs_prop {
    int n = 0;
    double *dev = nullptr;   // all table arrays, packed
    int *buckets = nullptr;  // bucket indexes of LIST_DH and LIST_H
    double *recs = nullptr;  // PropDev::recd and rech
    int *err = nullptr;
    PropDev p;
    hipStream_t stream = nullptr;
    std::string errmsg;
};

struct chs_eng {
    int L = 0, M = 0, N = 0;
    hipStream_t stream = nullptr;
    chs_prop *prop = nullptr;
    double *buf = nullptr;     // device: fields, history, coefficients
    int *err = nullptr;
    EngDev e{};
    double tgam[4], troh[4];
    size_t hs = 0;             // halo field size
    double *d0 = nullptr, *m0 = nullptr;   // DENSITY0, Viscousity0 (swapped with e.d, e.m each substep)
    double *gown = nullptr;                 // own G (e.g1.. may point at a momentum context's G)
    int rhs_lds = 1;                        // energy RHS from LDS-staged planes (k_eng_rhs_lds)
    int rhs_kc = 32;                        // its z-planes per block
    int rhs_nxcd = 8;                       // XCDs its tiles are grouped for
    std::string errmsg;
};

namespace {

int eng_fail(chs_eng *e, int code, const std::string &msg)
{
    if (e) e->errmsg = msg;
    g_prop_err = msg;
    return code;
}

// Reads and clears the table-range flags of the last launches (the reference STOPs there).
int eng_check(chs_eng *e, const char *what)
{
    int herr = 0;
    hipError_t he = hipGetLastError();
    if (he == hipSuccess) he = hipMemcpyAsync(&herr, e->err, sizeof(int), hipMemcpyDeviceToHost, e->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(e->stream);
    if (he != hipSuccess) return eng_fail(e, CHS_E_HIP, std::string(what) + ": " + hipGetErrorString(he));
    if (herr) {
        (void)hipMemsetAsync(e->err, 0, sizeof(int), e->stream);
        return eng_fail(e, CHS_E_RANGE, std::string(what) + ": DH outside the property table (flags " +
                                           std::to_string(herr) + ")");
    }
    return CHS_OK;
}

void eng_own_g(chs_eng *e)
{
    e->e.g1 = e->gown;
    e->e.g2 = e->gown + e->hs;
    e->e.g3 = e->gown + 2 * e->hs;
}

void eng_props_and_walls(chs_eng *e, int j0, int j1)
{
    const dim3 blk(64, 4), grd((e->L + 63) / 64, (j1 - j0 + 4) / 4, (e->N + 1) / 2);
    // without bucket indexes the bisections run on LDS copies of the two searched tables;
    // with them, a search touches two or three entries, which the caches serve
    const size_t sh = 2 * sizeof(double) * (size_t)e->prop->p.n;
    const int lds = sh <= 64 * 1024 && (e->prop->p.bdh.nb == 0 || e->prop->p.bh.nb == 0);
    hipLaunchKernelGGL(k_eng_props, grd, blk, lds ? sh : 0, e->stream, e->e, e->prop->p, j0, j1, e->err, lds);
    hipLaunchKernelGGL(k_eng_walls, dim3((e->L + 255) / 256, e->N), dim3(256), 0, e->stream, e->e, e->prop->p,
                       e->err);
}

}  // namespace

extern "C" {

int chs_eng_create(const chs_coeffs *c, const chs_prop_table *table, const chs_eng_params *p, chs_eng **out)
{
    if (!c || !table || !p || !out || c->L < 1 || c->M < 3 || c->N < 1 || !p->wallflux ||
        (p->wall_type[0] != 1 && p->wall_type[0] != 2) || (p->wall_type[1] != 1 && p->wall_type[1] != 2) ||
        p->top_face < 0 || p->top_face > c->M)
        return eng_fail(nullptr, CHS_E_ARG, "chs_eng_create: bad argument (wall types 1 or 2, top_face 0..M)");
    *out = nullptr;
    chs_eng *e = new chs_eng();
    const int L = c->L, M = c->M, N = c->N;
    e->L = L; e->M = M; e->N = N;
    e->hs = (size_t)N * (M + 2) * L;
    const size_t cs = (size_t)N * M * L;
    int rc = chs_prop_create(table, &e->prop);
    if (rc) {
        delete e;
        return rc;
    }
    // layout: 9 halo fields (dh h t d m k cp d0 m0), 3 G, rhs0, wallflux [2][N][L],
    // dyfi/dyci/wff/wfb (M + 2 each, reference index J)
    const size_t nb = 12 * e->hs + cs + 2 * (size_t)N * L + 4 * (size_t)(M + 2);
    hipError_t he = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (he == hipSuccess) he = hipMalloc(&e->buf, sizeof(double) * nb);
    if (he == hipSuccess) he = hipMalloc(&e->err, sizeof(int));
    if (he == hipSuccess) he = hipMemsetAsync(e->buf, 0, sizeof(double) * nb, e->stream);
    if (he == hipSuccess) he = hipMemsetAsync(e->err, 0, sizeof(int), e->stream);
    std::vector<double> geo(4 * (size_t)(M + 2), 0.0);
    for (int J = 1; J <= M; ++J) geo[J] = c->dyfi[J - 1];
    for (int J = 1; J <= M + 1; ++J) {
        geo[(M + 2) + J] = c->dyci[J - 1];
        geo[2 * (M + 2) + J] = c->wff[J - 1];
        geo[3 * (M + 2) + J] = c->wfb[J - 1];
    }
    double *f = e->buf;
    double *wfd = f + 12 * e->hs + cs, *gd = wfd + 2 * (size_t)N * L;
    if (he == hipSuccess)
        he = hipMemcpyAsync(wfd, p->wallflux, sizeof(double) * 2 * N * L, hipMemcpyHostToDevice, e->stream);
    if (he == hipSuccess)
        he = hipMemcpyAsync(gd, geo.data(), sizeof(double) * geo.size(), hipMemcpyHostToDevice, e->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(e->stream);
    if (he != hipSuccess) {
        chs_eng_destroy(e);
        return eng_fail(nullptr, CHS_E_HIP, std::string("chs_eng_create: ") + hipGetErrorString(he));
    }
    EngDev &d = e->e;
    d.L = L; d.M = M; d.N = N;
    d.dxi = c->dxi; d.dzi = c->dzi; d.dxqi = c->dxqi; d.dzqi = c->dzqi;
    d.dyfi = gd; d.dyci = gd + (M + 2); d.wff = gd + 2 * (M + 2); d.wfb = gd + 3 * (M + 2);
    double **fl[7] = {&d.dh, &d.h, &d.t, &d.d, &d.m, &d.k, &d.cp};
    for (int a = 0; a < 7; ++a) *fl[a] = f + a * e->hs;
    e->d0 = f + 7 * e->hs; e->m0 = f + 8 * e->hs;
    e->gown = f + 9 * e->hs;
    d.g1 = e->gown; d.g2 = e->gown + e->hs; d.g3 = e->gown + 2 * e->hs;
    d.rhs0 = f + 12 * e->hs;
    d.wallflux = wfd;
    d.wall_type0 = p->wall_type[0]; d.wall_type1 = p->wall_type[1];
    d.top_face = p->top_face > 0 ? p->top_face : M;
    d.cthecd = p->cthecd; d.dhmin = p->dh_min; d.dhmax = p->dh_max;
    for (int w = 0; w < 2; ++w)
        for (int j = 0; j < 5; ++j) d.ws[w][j] = p->wall_state[w][j];
    for (int s = 0; s < 4; ++s) { e->tgam[s] = c->tgam[s]; e->troh[s] = c->troh[s]; }
    *out = e;
    return CHS_OK;
}

void chs_eng_destroy(chs_eng *e)
{
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    if (e->buf) (void)hipFree(e->buf);
    if (e->err) (void)hipFree(e->err);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    if (e->prop) chs_prop_destroy(e->prop);
    delete e;
}

const char *chs_eng_last_error(const chs_eng *e) { return e ? e->errmsg.c_str() : g_prop_err.c_str(); }

int chs_eng_upload(chs_eng *e, const double *dh, const double *g, const double *rhs0, int init_props)
{
    if (!e || !dh || !g) return eng_fail(e, CHS_E_ARG, "chs_eng_upload: null argument");
    const size_t hsb = sizeof(double) * e->hs, cs = (size_t)e->N * e->M * e->L;
    eng_own_g(e);
    hipError_t he = hipMemcpyAsync(e->e.dh, dh, hsb, hipMemcpyHostToDevice, e->stream);
    if (he == hipSuccess) he = hipMemcpyAsync((double *)e->e.g1, g, 3 * hsb, hipMemcpyHostToDevice, e->stream);
    if (he == hipSuccess)
        he = rhs0 ? hipMemcpyAsync(e->e.rhs0, rhs0, sizeof(double) * cs, hipMemcpyHostToDevice, e->stream)
                  : hipMemsetAsync(e->e.rhs0, 0, sizeof(double) * cs, e->stream);
    if (he != hipSuccess) return eng_fail(e, CHS_E_HIP, std::string("chs_eng_upload: ") + hipGetErrorString(he));
    if (init_props) eng_props_and_walls(e, 0, e->M + 1);
    return eng_check(e, "chs_eng_upload");
}

int chs_eng_set_mass_flux(chs_eng *e, const double *g_dev)
{
    if (!e || !g_dev) return eng_fail(e, CHS_E_ARG, "chs_eng_set_mass_flux: null argument");
    eng_own_g(e);
    hipError_t he = hipMemcpyAsync((double *)e->e.g1, g_dev, 3 * sizeof(double) * e->hs, hipMemcpyDeviceToDevice,
                                   e->stream);
    if (he != hipSuccess) return eng_fail(e, CHS_E_HIP, std::string("chs_eng_set_mass_flux: ") + hipGetErrorString(he));
    return CHS_OK;
}

int chs_substep_eng(chs_eng *e, int ns, double dt)
{
    if (!e || ns < 1 || ns > 3 || !(dt > 0)) return eng_fail(e, CHS_E_ARG, "chs_substep_eng: bad argument");
    // Viscousity0 and DENSITY0 (SOLVERRK3_ENG_IO.f90:23-27), for the momentum substep: the
    // current arrays become the saved ones and the property update below rewrites every row of
    // D and M (rows 1..M here, the ghost rows in the wall update), so no copy is needed.  DH0
    // (:27) is never read by the reference and is not kept.
    std::swap(e->e.d, e->d0);
    std::swap(e->e.m, e->m0);
    // the LDS kernel loads column pairs as 16-byte words: the staged fields 16-byte aligned
    auto a16 = [](const void *q) { return ((uintptr_t)q & 15) == 0; };
    if (e->rhs_lds && e->L % 2 == 0 && a16(e->e.h) && a16(e->e.k) && a16(e->e.t) && a16(e->e.g1) && a16(e->e.g2) &&
        (e->hs % 2) == 0) {
        const int KC = e->rhs_kc;
        const dim3 grd((e->L + 63) / 64, (e->M + 3) / 4, (e->N + KC - 1) / KC);
        hipLaunchKernelGGL(k_eng_rhs_lds<4>, grd, dim3(64, 4), 0, e->stream, e->e, e->tgam[ns] * dt, e->troh[ns] * dt,
                           KC, e->rhs_nxcd);
    } else {
        constexpr int KC = 8;
        const dim3 blk(64, 4), grd((e->L + 63) / 64, (e->M + 3) / 4, (e->N + KC - 1) / KC);
        hipLaunchKernelGGL(k_eng_rhs_update, grd, blk, 0, e->stream, e->e, e->tgam[ns] * dt, e->troh[ns] * dt, KC);
    }
    eng_props_and_walls(e, 1, e->M);
    return eng_check(e, "chs_substep_eng");
}

int chs_eng_set_tuning(chs_eng *e, const char *key, int value)
{
    if (!e || !key) return eng_fail(e, CHS_E_ARG, "chs_eng_set_tuning: bad argument");
    const std::string k(key);
    if (k == "rhs_lds" && (value == 0 || value == 1)) e->rhs_lds = value;
    else if (k == "rhs_kc" && value >= 1 && value <= 1024) e->rhs_kc = value;
    else if (k == "rhs_nxcd" && value >= 1 && value <= 64) e->rhs_nxcd = value;
    else return eng_fail(e, CHS_E_ARG, "chs_eng_set_tuning: unknown key or value out of range (" + k + ")");
    return CHS_OK;
}

int chs_eng_download(chs_eng *e, double *dh, double *h, double *t, double *d, double *m, double *k, double *cp,
                     double *rhs0)
{
    if (!e) return eng_fail(e, CHS_E_ARG, "chs_eng_download: null engine");
    const size_t hsb = sizeof(double) * e->hs, cs = (size_t)e->N * e->M * e->L;
    double *dst[7] = {dh, h, t, d, m, k, cp};
    const double *src[7] = {e->e.dh, e->e.h, e->e.t, e->e.d, e->e.m, e->e.k, e->e.cp};
    hipError_t he = hipSuccess;
    for (int a = 0; a < 7 && he == hipSuccess; ++a)
        if (dst[a]) he = hipMemcpyAsync(dst[a], src[a], hsb, hipMemcpyDeviceToHost, e->stream);
    if (he == hipSuccess && rhs0)
        he = hipMemcpyAsync(rhs0, e->e.rhs0, sizeof(double) * cs, hipMemcpyDeviceToHost, e->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(e->stream);
    if (he != hipSuccess) return eng_fail(e, CHS_E_HIP, std::string("chs_eng_download: ") + hipGetErrorString(he));
    return CHS_OK;
}

int chs_prop_table_setup(int n, const double *rows, double T0, double *list, double *spl, int *idx,
                         double *ref)
{
#pragma clang fp contract(off)
    if (n < 4 || !rows || !list || !spl || !idx || !ref) {
        g_prop_err = "chs_prop_table_setup: bad argument";
        return CHS_E_ARG;
    }
    std::vector<double> H(n), T(n), D(n), M(n), K(n), CP(n), B(n);
    for (int i = 0; i < n; ++i) {
        H[i] = rows[8 * i + 1]; T[i] = rows[8 * i + 2]; D[i] = rows[8 * i + 3]; M[i] = rows[8 * i + 4];
        K[i] = rows[8 * i + 5]; CP[i] = rows[8 * i + 6]; B[i] = rows[8 * i + 7];
    }
    // sort by T, selection with the first minimum (thermo_properties_table_searching.f90:1267-1293)
    for (int i = 0; i < n; ++i) {
        int k = i;
        for (int j = i + 1; j < n; ++j)
            if (T[j] < T[k]) k = j;
        std::swap(T[i], T[k]); std::swap(D[i], D[k]); std::swap(H[i], H[k]); std::swap(M[i], M[k]);
        std::swap(K[i], K[k]); std::swap(B[i], B[k]); std::swap(CP[i], CP[k]);
    }
    // reference state by bisection + linear interpolation at T0 (BiSection_LIST_T, :808-839;
    // :1297-1306)
    int is = 1, ie = n;
    while (ie - is > 1) {
        const int im = is + (ie - is) / 2;
        const double vs = T[is - 1] - T0, vm = T[im - 1] - T0;
        if (vs * vm > 0.0) is = im;
        else ie = im;
    }
    const double sfs = (T[ie - 1] - T0) / (T[ie - 1] - T[is - 1]);
    const double sfe = 1.0 - sfs;
    const double H0 = sfe * H[ie - 1] + sfs * H[is - 1];
    const double D0 = sfe * D[ie - 1] + sfs * D[is - 1];
    const double K0 = sfe * K[ie - 1] + sfs * K[is - 1];
    const double M0 = sfe * M[ie - 1] + sfs * M[is - 1];
    const double CP0 = sfe * CP[ie - 1] + sfs * CP[is - 1];
    const double B0 = sfe * B[ie - 1] + sfs * B[is - 1];
    std::vector<double> DH(n);
    for (int i = 0; i < n; ++i) {   // :1311-1320
        H[i] = (H[i] - H0) / T0 / CP0;
        T[i] = T[i] / T0;
        D[i] = D[i] / D0;
        M[i] = M[i] / M0;
        K[i] = K[i] / K0;
        CP[i] = CP[i] / CP0;
        B[i] = B[i] / B0;
        DH[i] = D[i] * H[i];
    }
    idx[0] = first_min(DH);
    idx[1] = first_max(DH);
    idx[2] = first_min(H);
    idx[3] = first_max(H);
    const std::vector<double> *cols[8] = {&DH, &H, &T, &D, &M, &K, &CP, &B};
    for (int c = 0; c < 8; ++c)
        for (int i = 0; i < n; ++i) list[(size_t)c * n + i] = (*cols[c])[i];
    // SPLINE_COEFFICIENTS_FOR_DH / _FOR_H (:977-1129): DH->H, then H->T, D, M, K, Cp, B
    double *s = spl;
    cubic_spline(n, DH.data(), H.data(), s, s + n, s + 2 * n);
    const std::vector<double> *ys[6] = {&T, &D, &M, &K, &CP, &B};
    for (int c = 0; c < 6; ++c) {
        double *o = spl + (size_t)(c + 1) * 3 * n;
        cubic_spline(n, H.data(), ys[c]->data(), o, o + n, o + 2 * n);
    }
    ref[0] = T0; ref[1] = H0; ref[2] = D0; ref[3] = M0; ref[4] = K0; ref[5] = CP0; ref[6] = B0;
    return CHS_OK;
}

int chs_prop_create(const chs_prop_table *t, chs_prop **out)
{
    if (!t || !out || t->n < 4) {
        g_prop_err = "chs_prop_create: bad argument";
        return CHS_E_ARG;
    }
    *out = nullptr;
    chs_prop *p = new chs_prop();
    const int n = t->n;
    p->n = n;
    const double *src[29] = {t->dh, t->h, t->t, t->d, t->m, t->k, t->cp, t->b,
                             t->dhh_b, t->dhh_c, t->dhh_d, t->ht_b, t->ht_c, t->ht_d,
                             t->hd_b, t->hd_c, t->hd_d, t->hm_b, t->hm_c, t->hm_d,
                             t->hk_b, t->hk_c, t->hk_d, t->hcp_b, t->hcp_c, t->hcp_d,
                             t->hb_b, t->hb_c, t->hb_d};
    std::vector<double> pack((size_t)29 * n);
    for (int a = 0; a < 29; ++a) {
        if (!src[a]) {
            delete p;
            g_prop_err = "chs_prop_create: null table array";
            return CHS_E_ARG;
        }
        for (int i = 0; i < n; ++i) pack[(size_t)a * n + i] = src[a][i];
    }
    hipError_t e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&p->dev, sizeof(double) * pack.size());
    if (e == hipSuccess) e = hipMalloc(&p->err, sizeof(int));
    if (e == hipSuccess) e = hipMemcpyAsync(p->dev, pack.data(), sizeof(double) * pack.size(),
                                            hipMemcpyHostToDevice, p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    if (e != hipSuccess) {
        g_prop_err = std::string("chs_prop_create: ") + hipGetErrorString(e);
        chs_prop_destroy(p);
        return CHS_E_HIP;
    }
    auto A = [&](int a) { return (const double *)(p->dev + (size_t)a * n); };
    p->p.n = n;
    p->p.dhh = {A(0), A(1), A(8), A(9), A(10), t->imin_dh, t->imax_dh};
    const int ys[6] = {2, 3, 4, 5, 6, 7};   // T, D, M, K, Cp, B
    SplineDev *dst[6] = {&p->p.ht, &p->p.hd, &p->p.hm, &p->p.hk, &p->p.hcp, &p->p.hb};
    for (int c = 0; c < 6; ++c) *dst[c] = {A(1), A(ys[c]), A(11 + 3 * c), A(12 + 3 * c), A(13 + 3 * c), t->imin_h, t->imax_h};
    // bucket indexes (2 n buckets) for the two searched tables where they are non-decreasing
    const int nb = 2 * n;
    std::vector<int> bk((size_t)4 * nb);
    BucketDev *bd[2] = {&p->p.bdh, &p->p.bh};
    const double *xs[2] = {t->dh, t->h};
    for (int a = 0; a < 2; ++a) {
        *bd[a] = {nullptr, 0.0, 0.0, 0};
        const double *x = xs[a];
        bool mono = x[n - 1] > x[0];
        for (int i = 1; i < n && mono; ++i) mono = x[i] >= x[i - 1];
        if (!mono) continue;
        auto bis = [&](double xv) {
            int I = 1, J = n + 1;
            while (J > I + 1) {
                const int K = (I + J) / 2;
                if (xv < x[K - 1]) J = K;
                else I = K;
            }
            return I;
        };
        const double x0 = x[0], w = (x[n - 1] - x[0]) / (double)nb;
        int *r = bk.data() + (size_t)a * 2 * nb;
        for (int b = 0; b < nb; ++b) {
            r[2 * b] = b == 0 ? 1 : std::max(1, bis(x0 + (double)b * w) - 1);
            r[2 * b + 1] = b == nb - 1 ? n : std::min(n, bis(x0 + (double)(b + 1) * w) + 1);
        }
        *bd[a] = {nullptr, x0, (double)nb / (x[n - 1] - x[0]), nb};
    }
    e = hipMalloc(&p->buckets, sizeof(int) * bk.size());
    if (e == hipSuccess)
        e = hipMemcpyAsync(p->buckets, bk.data(), sizeof(int) * bk.size(), hipMemcpyHostToDevice, p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    if (e != hipSuccess) {
        g_prop_err = std::string("chs_prop_create: ") + hipGetErrorString(e);
        chs_prop_destroy(p);
        return CHS_E_HIP;
    }
    p->p.bdh.lohi = p->buckets;
    p->p.bh.lohi = p->buckets + 2 * (size_t)nb;
    // evaluation records (PropDev::recd, rech)
    std::vector<double> rec((size_t)32 * n, 0.0);
    for (int i = 0; i < n; ++i) {
        double *rd = rec.data() + (size_t)8 * i, *rh = rec.data() + (size_t)8 * n + (size_t)24 * i;
        const double vd[5] = {t->dh[i], t->h[i], t->dhh_b[i], t->dhh_c[i], t->dhh_d[i]};
        for (int q = 0; q < 5; ++q) rd[q] = vd[q];
        rh[0] = t->h[i];
        const double *hs[5][4] = {{t->t, t->ht_b, t->ht_c, t->ht_d}, {t->m, t->hm_b, t->hm_c, t->hm_d},
                                  {t->k, t->hk_b, t->hk_c, t->hk_d}, {t->d, t->hd_b, t->hd_c, t->hd_d},
                                  {t->cp, t->hcp_b, t->hcp_c, t->hcp_d}};
        for (int sp = 0; sp < 5; ++sp)
            for (int q = 0; q < 4; ++q) rh[1 + 4 * sp + q] = hs[sp][q][i];
    }
    e = hipMalloc(&p->recs, sizeof(double) * rec.size());
    if (e == hipSuccess)
        e = hipMemcpyAsync(p->recs, rec.data(), sizeof(double) * rec.size(), hipMemcpyHostToDevice, p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    if (e != hipSuccess) {
        g_prop_err = std::string("chs_prop_create: ") + hipGetErrorString(e);
        chs_prop_destroy(p);
        return CHS_E_HIP;
    }
    p->p.recd = p->recs;
    p->p.rech = p->recs + (size_t)8 * n;
    *out = p;
    return CHS_OK;
}

void chs_prop_destroy(chs_prop *p)
{
    if (!p) return;
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    if (p->dev) (void)hipFree(p->dev);
    if (p->buckets) (void)hipFree(p->buckets);
    if (p->recs) (void)hipFree(p->recs);
    if (p->err) (void)hipFree(p->err);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

const char *chs_prop_last_error(const chs_prop *p) { return p ? p->errmsg.c_str() : g_prop_err.c_str(); }

int chs_prop_update_from_dh(chs_prop *p, long n, const double *dh, double *out)
{
    if (!p || !dh || !out || n < 0) {
        g_prop_err = "chs_prop_update_from_dh: bad argument";
        return CHS_E_ARG;
    }
    if (n == 0) return CHS_OK;
    double *d_in = nullptr, *d_out = nullptr;
    int herr = 0;
    hipError_t e = hipMalloc(&d_in, sizeof(double) * n);
    if (e == hipSuccess) e = hipMalloc(&d_out, sizeof(double) * 7 * n);
    if (e == hipSuccess) e = hipMemcpyAsync(d_in, dh, sizeof(double) * n, hipMemcpyHostToDevice, p->stream);
    if (e == hipSuccess) e = hipMemsetAsync(p->err, 0, sizeof(int), p->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_prop_from_dh, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, p->stream, p->p, n,
                           d_in, d_out, p->err);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, sizeof(double) * 7 * n, hipMemcpyDeviceToHost, p->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&herr, p->err, sizeof(int), hipMemcpyDeviceToHost, p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    if (d_in) (void)hipFree(d_in);
    if (d_out) (void)hipFree(d_out);
    if (e != hipSuccess) {
        p->errmsg = std::string("chs_prop_update_from_dh: ") + hipGetErrorString(e);
        return CHS_E_HIP;
    }
    if (herr) {
        p->errmsg = "chs_prop_update_from_dh: value outside the property table (flags " + std::to_string(herr) + ")";
        return CHS_E_RANGE;
    }
    return CHS_OK;
}

}  // extern "C"

// Device fields of an energy context that the coupled heated-channel substep hands to the
// momentum substep: its stream, DENSITY0, Viscousity0, Viscousity (the current viscosity,
// read by the Crank-Nicolson momentum terms) and the mass-flux array G.
__attribute__((visibility("hidden"))) void chs_internal_eng_fields(chs_eng *e, hipStream_t *s, double **d0,
                                                                   double **m0, double **m, double **g)
{
    *s = e->stream;
    *d0 = e->d0;
    *m0 = e->m0;
    *m = e->e.m;
    *g = const_cast<double *>(e->e.g1);
}

// The current DENSITY of an energy context (CFL_io reads it).
__attribute__((visibility("hidden"))) double *chs_internal_eng_density(chs_eng *e) { return e->e.d; }

// Point an energy context's G at a momentum context's G array (3 halo fields) instead of
// copying it (NULL: back to the context's own array, as chs_eng_upload and
// chs_eng_set_mass_flux also do).
__attribute__((visibility("hidden"))) void chs_internal_eng_set_g(chs_eng *e, const double *g)
{
    if (!g) g = e->gown;
    e->e.g1 = g;
    e->e.g2 = g + e->hs;
    e->e.g3 = g + 2 * e->hs;
}
