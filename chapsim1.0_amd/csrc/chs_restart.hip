// Instantaneous-field files in the reference's restart format (host code), for the
// turbulence-generator and the main domain: WRT_INSTANT_VARS_TG (processing/WRT_VARS_TG.f90:
// 12-104) and WRT_INSTANT_VARS_IO (processing/WRT_VARS_IO.f90:16-227) write, per
// variable U, V, W, P, the file <dir>/DNS_perixz_INSTANT_T<PhyTIME as ES15.9>_<var>.D holding
// int32 NCL1, NCL2, NCL3, ITERG, then float64 PhyTIME, REN, DT, then the NCL1 x NCL2 x NCL3
// interior values in Fortran order (MPI native representation, little-endian here);
// ReStart_INSTANT_VARS_TG (processing/RESTART_tg.f90:11-56) with READ_3D_VARS
// (processing/READ_INTP_3D_VARS.f90:20-103) reads them back.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/chapsim_hip.h"

namespace {

thread_local std::string g_rst_err;

int rst_fail(int code, const std::string &msg)
{
    g_rst_err = msg;
    return code;
}

std::string rst_name(const char *dir, double t, char var)
{
    char tm[64];
    std::snprintf(tm, sizeof tm, "%.9E", t);   // Fortran ES15.9 for |t| < 1e100
    std::string d(dir ? dir : "");
    if (!d.empty() && d.back() != '/') d += '/';
    return d + "DNS_perixz_INSTANT_T" + tm + "_" + var + ".D";
}

const char kVars[4] = {'U', 'V', 'W', 'P'};

// One variable file of the format above from the interior rows of a halo array f.
int write_var(const char *who, const char *dir, char var, int L, int M, int N, int iterg, double phytime, double ren,
              double dt, const double *f, std::vector<double> &buf)
{
    for (int k = 0; k < N; ++k)
        for (int J = 1; J <= M; ++J)
            for (int i = 0; i < L; ++i) buf[((size_t)k * M + (J - 1)) * L + i] = f[((size_t)k * (M + 2) + J) * L + i];
    const std::string fn = rst_name(dir, phytime, var);
    FILE *fp = std::fopen(fn.c_str(), "wb");
    if (!fp) return rst_fail(CHS_E_ARG, std::string(who) + ": cannot open " + fn);
    const int32_t hdr[4] = {L, M, N, iterg};
    const double rl[3] = {phytime, ren, dt};
    const bool ok = std::fwrite(hdr, sizeof hdr, 1, fp) == 1 && std::fwrite(rl, sizeof rl, 1, fp) == 1 &&
                    std::fwrite(buf.data(), sizeof(double), buf.size(), fp) == buf.size();
    if (std::fclose(fp) != 0 || !ok) return rst_fail(CHS_E_ARG, std::string(who) + ": write failed: " + fn);
    return CHS_OK;
}

// One variable file into the interior rows of a halo array f (READ_3D_VARS,
// processing/READ_INTP_3D_VARS.f90:20-103); the header's ITERG, PhyTIME to iterg, phytime.
int read_var(const char *who, const char *dir, double tfm, char var, int L, int M, int N, int *iterg, double *phytime,
             double *f, std::vector<double> &buf)
{
    const std::string fn = rst_name(dir, tfm, var);
    FILE *fp = std::fopen(fn.c_str(), "rb");
    if (!fp) return rst_fail(CHS_E_ARG, std::string(who) + ": file " + fn + " does not exist");
    int32_t hdr[4];
    double rl[3];
    const bool ok = std::fread(hdr, sizeof hdr, 1, fp) == 1 && std::fread(rl, sizeof rl, 1, fp) == 1 &&
                    std::fread(buf.data(), sizeof(double), buf.size(), fp) == buf.size();
    std::fclose(fp);
    if (!ok) return rst_fail(CHS_E_ARG, std::string(who) + ": short file " + fn);
    if (hdr[0] != L || hdr[1] != M || hdr[2] != N)
        return rst_fail(CHS_E_ARG, std::string(who) + ": " + fn + " holds " + std::to_string(hdr[0]) + " x " +
                                       std::to_string(hdr[1]) + " x " + std::to_string(hdr[2]));
    if (iterg) *iterg = hdr[3];
    if (phytime) *phytime = rl[0];
    for (int k = 0; k < N; ++k)
        for (int J = 1; J <= M; ++J)
            for (int i = 0; i < L; ++i) f[((size_t)k * (M + 2) + J) * L + i] = buf[((size_t)k * M + (J - 1)) * L + i];
    return CHS_OK;
}

// Wall rows after a restart read: velocity (or mass flux) ghost rows and v on J = 1 zero,
// pressure ghosts copy the wall cells (BC_WALL_Q_tg / BC_WALL_G_io, BC_WALL_PR_TG / _io).
void restart_walls(int L, int M, int N, double *q, double *pr)
{
    const size_t hs = (size_t)N * (M + 2) * L;
    for (int k = 0; k < N; ++k)
        for (int i = 0; i < L; ++i) {
            for (int c = 0; c < 3; ++c) {
                q[c * hs + ((size_t)k * (M + 2)) * L + i] = 0.0;
                q[c * hs + ((size_t)k * (M + 2) + M + 1) * L + i] = 0.0;
            }
            q[hs + ((size_t)k * (M + 2) + 1) * L + i] = 0.0;
            pr[((size_t)k * (M + 2)) * L + i] = pr[((size_t)k * (M + 2) + 1) * L + i];
            pr[((size_t)k * (M + 2) + M + 1) * L + i] = pr[((size_t)k * (M + 2) + M) * L + i];
        }
}

}  // namespace

extern "C" {

int chs_restart_write_tg(const char *dir, int L, int M, int N, int iterg, double phytime, double ren, double dt,
                         const double *q, const double *pr)
{
    if (L < 1 || M < 1 || N < 1 || !q || !pr) return rst_fail(CHS_E_ARG, "chs_restart_write_tg: bad argument");
    const size_t hs = (size_t)N * (M + 2) * L;
    std::vector<double> buf((size_t)L * M * N);
    for (int v = 0; v < 4; ++v)
        if (int rc = write_var("chs_restart_write_tg", dir, kVars[v], L, M, N, iterg, phytime, ren, dt,
                               v < 3 ? q + v * hs : pr, buf))
            return rc;
    return CHS_OK;
}

int chs_restart_read_tg(const char *dir, double tfm, int L, int M, int N, int *iterg, double *phytime, double *q,
                        double *pr)
{
    if (L < 1 || M < 1 || N < 1 || !q || !pr) return rst_fail(CHS_E_ARG, "chs_restart_read_tg: bad argument");
    const size_t hs = (size_t)N * (M + 2) * L;
    std::vector<double> buf((size_t)L * M * N);
    for (int v = 0; v < 4; ++v)
        if (int rc = read_var("chs_restart_read_tg", dir, tfm, kVars[v], L, M, N, v == 0 ? iterg : nullptr,
                              v == 0 ? phytime : nullptr, v < 3 ? q + v * hs : pr, buf))
            return rc;
    // walls as FLOWSTART applies them after the read (FLOWSTART.f90:37-40)
    restart_walls(L, M, N, q, pr);
    return CHS_OK;
}

int chs_restart_write_io(const char *dir, int L, int M, int N, int iterg, double phytime, double ren, double dt,
                         const double *q, const double *pr, const double *t, const double *d, const double *h,
                         const double *m, const double *k)
{
    const bool th = t || d || h || m || k;
    if (L < 1 || M < 1 || N < 1 || !q || !pr || (th && !(t && d && h && m && k)))
        return rst_fail(CHS_E_ARG, "chs_restart_write_io: bad argument (thermal fields: all five or none)");
    const size_t hs = (size_t)N * (M + 2) * L;
    std::vector<double> buf((size_t)L * M * N);
    // WRT_INSTANT_VARS_FLOW_IO (processing/WRT_VARS_IO.f90:24-123): Q_io, PR_io
    for (int v = 0; v < 4; ++v)
        if (int rc = write_var("chs_restart_write_io", dir, kVars[v], L, M, N, iterg, phytime, ren, dt,
                               v < 3 ? q + v * hs : pr, buf))
            return rc;
    if (!th) return CHS_OK;
    // WRT_INSTANT_VARS_THERMAL_IO (:126-227): TEMPERATURE, DENSITY, ENTHALPY, Viscousity, THERMCONDT
    const char tv[5] = {'T', 'D', 'H', 'M', 'K'};
    const double *tf[5] = {t, d, h, m, k};
    for (int v = 0; v < 5; ++v)
        if (int rc = write_var("chs_restart_write_io", dir, tv[v], L, M, N, iterg, phytime, ren, dt, tf[v], buf))
            return rc;
    return CHS_OK;
}

int chs_restart_read_io(const char *dir, double tfm, int L, int M, int N, int ini_field_type, int *iterg,
                        double *phytime, double *flow, double *pr, double *t, double *d, double *h, double *dh)
{
    const bool th = t || d || h || dh;
    if (L < 1 || M < 1 || N < 1 || !flow || !pr || (ini_field_type != 0 && ini_field_type != 1) ||
        (th && !(t && d && h && dh)) || (th && ini_field_type != 0))
        return rst_fail(CHS_E_ARG, "chs_restart_read_io: bad argument (type 0 or 1; thermal fields: all four, "
                                   "type 0 only)");
    const size_t hs = (size_t)N * (M + 2) * L;
    std::vector<double> buf((size_t)L * M * N);
    // ReStart_INSTANT_VARS_io (processing/RESTART_io.f90:16-92): U, V, W, P into G_io (type 0)
    // or Q_io (type 1), and PR_io
    for (int v = 0; v < 4; ++v)
        if (int rc = read_var("chs_restart_read_io", dir, tfm, kVars[v], L, M, N, v == 0 ? iterg : nullptr,
                              v == 0 ? phytime : nullptr, v < 3 ? flow + v * hs : pr, buf))
            return rc;
    restart_walls(L, M, N, flow, pr);   // BC_WALL_G_io / BC_WALL_PR_io (:147-157)
    if (!th) return CHS_OK;
    // thermal field (:95-131): TEMPERATURE, DENSITY, ENTHALPY, then DH = DENSITY * ENTHALPY
    const char tv[3] = {'T', 'D', 'H'};
    double *tf[3] = {t, d, h};
    for (int v = 0; v < 3; ++v)
        if (int rc = read_var("chs_restart_read_io", dir, tfm, tv[v], L, M, N, nullptr, nullptr, tf[v], buf))
            return rc;
    for (int kk = 0; kk < N; ++kk)
        for (int J = 1; J <= M; ++J)
            for (int i = 0; i < L; ++i) {
                const size_t o = ((size_t)kk * (M + 2) + J) * L + i;
                dh[o] = d[o] * h[o];
            }
    return CHS_OK;
}

const char *chs_restart_last_error(void) { return g_rst_err.c_str(); }

}  // extern "C"

// ---------------------------------------------------------------------------------------
// Interpolated start from a coarser field: INITIAL_INTERP_tg (processing/INTERPOLATION_INITIAL.f90:
// 189-300), INITIAL_INTERP_MESH (:51-186), INTP_3D_VARS (processing/READ_INTP_3D_VARS.f90:106-165).
// ---------------------------------------------------------------------------------------
namespace {

// SPLINE with both end slopes given as 0 (the callers pass DP1 = DPN = 0,
// READ_INTP_3D_VARS.f90:341-386): second derivatives y2 of the interpolating cubic spline.
void spline0(int n, const double *x, const double *y, double *y2, double *u)
{
    y2[0] = -0.5;
    u[0] = (3.0 / (x[1] - x[0])) * ((y[1] - y[0]) / (x[1] - x[0]) - 0.0);
    for (int i = 1; i < n - 1; ++i) {
        const double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
        const double p = sig * y2[i - 1] + 2.0;
        y2[i] = (sig - 1.0) / p;
        u[i] = (6.0 * ((y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1])) /
                    (x[i + 1] - x[i - 1]) -
                sig * u[i - 1]) /
               p;
    }
    const double qn = 0.5, un = (3.0 / (x[n - 1] - x[n - 2])) * (0.0 - (y[n - 1] - y[n - 2]) / (x[n - 1] - x[n - 2]));
    y2[n - 1] = (un - qn * u[n - 2]) / (qn * y2[n - 2] + 1.0);
    for (int k = n - 2; k >= 0; --k) y2[k] = y2[k] * y2[k + 1] + u[k];
}

// SPLINT (READ_INTP_3D_VARS.f90:389-432): bisection, then the cubic of the bracketing interval
// (extrapolated beyond the end points).
double splint(int n, const double *xa, const double *ya, const double *y2a, double x)
{
    int klo = 1, khi = n;
    while (khi - klo > 1) {
        const int k = (khi + klo) / 2;
        if (xa[k - 1] > x) khi = k;
        else klo = k;
    }
    const double h = xa[khi - 1] - xa[klo - 1];
    const double a = (xa[khi - 1] - x) / h, b = (x - xa[klo - 1]) / h;
    return a * ya[klo - 1] + b * ya[khi - 1] +
           ((a * a * a - a) * y2a[klo - 1] + (b * b * b - b) * y2a[khi - 1]) * (h * h) / 6.0;
}

// One 1-D interpolation along a line of n old values with stride so, onto nn new points with
// stride sn.
void interp_line(int n, const double *xo, const double *fo, long so, int nn, const double *xn, double *fn, long sn,
                 std::vector<double> &buf)
{
    buf.resize(4 * (size_t)n);
    double *f = buf.data(), *y2 = f + n, *u = y2 + n;
    for (int i = 0; i < n; ++i) f[i] = fo[(long)i * so];
    spline0(n, xo, f, y2, u);
    for (int i = 0; i < nn; ++i) fn[(long)i * sn] = splint(n, xo, f, y2, xn[i]);
}

}  // namespace

extern "C" int chs_interp_initial_tg(int Loo, int Mo, int Noo, double HXoo, double HZoo, const double *yndo,
                                     const double *uo, const double *vo, const double *wo, const double *po, int L,
                                     int M, int N, double HX, double HZ, double STR2, double *q, double *pr)
{
    if (Loo < 2 || Mo < 2 || Noo < 2 || L < 2 || M < 2 || N < 2 || !yndo || !uo || !vo || !wo || !po || !q || !pr ||
        !(HXoo > 0) || !(HZoo > 0) || !(HX > 0) || !(HZ > 0))
        return rst_fail(CHS_E_ARG, "chs_interp_initial_tg: bad argument");
    // INITIAL_INTERP_MESH: the given box repeated to cover the target box (:113-114)
    const int Lo = (int)std::ceil(HX / (HXoo / (double)Loo)), No = (int)std::ceil(HZ / (HZoo / (double)Noo));
    std::vector<double> xndo(Lo + 1), xcco(Lo), zndo(No + 1), zcco(No), ycco(Mo);
    for (int i = 0; i <= Lo; ++i) xndo[i] = (double)i * HX / (double)Lo;
    for (int k = 0; k <= No; ++k) zndo[k] = (double)k * HZ / (double)No;
    for (int i = 0; i < Lo; ++i) xcco[i] = 0.5 * (xndo[i] + xndo[i + 1]);
    for (int k = 0; k < No; ++k) zcco[k] = 0.5 * (zndo[k] + zndo[k + 1]);
    for (int j = 0; j < Mo; ++j) ycco[j] = 0.5 * (yndo[j] + yndo[j + 1]);
    // target mesh (CONSPARA: DX = HX / NCL1, DZ = HZ / NCL3; COORDJR.f90:44-55, :166-194)
    const double dx = HX / (double)L, dz = HZ / (double)N;
    std::vector<double> xnd(L + 1), xcc(L), znd(N + 1), zcc(N), ynd(M + 1), ycc(M), dyfi(M);
    for (int i = 0; i <= L; ++i) xnd[i] = (double)i * dx;
    for (int i = 0; i < L; ++i) xcc[i] = 0.5 * (xnd[i] + xnd[i + 1]);
    for (int k = 0; k <= N; ++k) znd[k] = (double)k * dz;
    for (int k = 0; k < N; ++k) zcc[k] = 0.5 * (znd[k] + znd[k + 1]);
    for (int j = 0; j <= M; ++j) {
        const double x2 = (double)j / (double)M;
        ynd[j] = std::tanh(STR2 * (x2 - 0.5)) / std::tanh(STR2 * 0.5);
    }
    for (int j = 0; j < M; ++j) {
        ycc[j] = (ynd[j] + ynd[j + 1]) * 0.5;
        dyfi[j] = 1.0 / (ynd[j + 1] - ynd[j]);
    }
    const size_t hs = (size_t)N * (M + 2) * L;
    const double *src[4] = {uo, vo, wo, po};
    std::vector<double> qo((size_t)Lo * Mo * No), qx((size_t)L * No), qxz((size_t)L * Mo * N), buf;
    for (int v = 0; v < 4; ++v) {
        // QO(I, J, K) = DUMMY(MOD(I, NCLOO1), J, MOD(K, NCLOO3)) (:221-231), layout [k][j][i]
        for (int k = 0; k < No; ++k)
            for (int j = 0; j < Mo; ++j)
                for (int i = 0; i < Lo; ++i)
                    qo[((size_t)k * Mo + j) * Lo + i] = src[v][((size_t)(k % Noo) * Mo + j) * Loo + (i % Loo)];
        const double *xo = v == 0 ? xndo.data() : xcco.data(), *xn = v == 0 ? xnd.data() : xcc.data();
        const double *zo = v == 2 ? zndo.data() : zcco.data(), *zn = v == 2 ? znd.data() : zcc.data();
        const double *yo = v == 1 ? yndo : ycco.data(), *yn = v == 1 ? ynd.data() : ycc.data();
        for (int j = 0; j < Mo; ++j) {
            for (int k = 0; k < No; ++k)   // x: INTRPRI (:172-206)
                interp_line(Lo, xo, &qo[((size_t)k * Mo + j) * Lo], 1, L, xn, &qx[(size_t)k * L], 1, buf);
            for (int i = 0; i < L; ++i)    // z: INTRPRK (:215-247)
                interp_line(No, zo, &qx[i], L, N, zn, &qxz[(size_t)j * L + i], (long)Mo * L, buf);
        }
        double *dst = v < 3 ? q + v * hs : pr;
        for (int k = 0; k < N; ++k)        // y: INTRPRJG (:290-337), into rows 1..M
            for (int i = 0; i < L; ++i)
                interp_line(Mo, yo, &qxz[(size_t)k * Mo * L + i], L, M, yn, &dst[((size_t)k * (M + 2) + 1) * L + i],
                            L, buf);
    }
    // u / bulk velocity (BULK_VELOCITY_TG, INITIAL_FL_FLD_TG.f90:142-166; Area_inlet
    // COORDJR.f90:368-373), then the walls of FLOWSTART.f90:37-40
    double um = 0.0;
    for (int j = 1; j <= M; ++j)
        for (int i = 0; i < L; ++i)
            for (int k = 0; k < N; ++k) um = um + q[((size_t)k * (M + 2) + j) * L + i] / dyfi[j - 1];
    um = um / (1.0 / dz);
    double area = 0.0;
    for (int j = 0; j < M; ++j)
        for (int k = 0; k < N; ++k) area = area + 1.0 / dyfi[j] / (1.0 / dz);
    um = um / (area * (double)L);
    for (int k = 0; k < N; ++k)
        for (int j = 1; j <= M; ++j)
            for (int i = 0; i < L; ++i) q[((size_t)k * (M + 2) + j) * L + i] = q[((size_t)k * (M + 2) + j) * L + i] / um;
    for (int k = 0; k < N; ++k)
        for (int i = 0; i < L; ++i) {
            for (int c = 0; c < 3; ++c) {
                q[c * hs + ((size_t)k * (M + 2)) * L + i] = 0.0;
                q[c * hs + ((size_t)k * (M + 2) + M + 1) * L + i] = 0.0;
            }
            q[hs + ((size_t)k * (M + 2) + 1) * L + i] = 0.0;
            pr[((size_t)k * (M + 2)) * L + i] = pr[((size_t)k * (M + 2) + 1) * L + i];
            pr[((size_t)k * (M + 2) + M + 1) * L + i] = pr[((size_t)k * (M + 2) + M) * L + i];
        }
    return CHS_OK;
}

// ---------------------------------------------------------------------------------------
// Random (or sine) start of the turbulence-generator channel: random_FL_FLD_TG
// (INITIAL_FL_FLD_TG.f90:16-104) with IniField_FLOW_tg (:174-244), INITIAL_MEAN_TG
// (processing/INITIAL_MEAN.f90:13-122), the laminar profile Vini (LAMPOISLPROF.f90:29-51) and
// BULK_VELOCITY_TG (INITIAL_FL_FLD_TG.f90:142-166).
// ---------------------------------------------------------------------------------------
namespace {

// splitmix64: a counter-based stand-in for the compiler's RANDOM_NUMBER, which no other
// compiler reproduces; the value at a point is a pure function of its key.
uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// rvec_random(-1, 1, 3, XRD) (processing/randomgen.f90): a + (b - a) r with r uniform in [0, 1)
void rvec3(uint64_t key, double *x)
{
    for (int c = 0; c < 3; ++c) {
        const double r = (double)(mix64(key * 3ull + (uint64_t)c) >> 11) * 0x1.0p-53;
        x[c] = -1.0 + 2.0 * r;
    }
}

}  // namespace

extern "C" int chs_random_field_tg(int L, int M, int N, double HX, double HZ, double STR2, int random_type,
                                   double vperg, double svperg, unsigned long long seed, double m_over_d, double *q,
                                   double *pr)
{
    if (L < 2 || M < 2 || N < 2 || !q || !pr || !(HX > 0) || !(HZ > 0) || random_type < 1 || random_type > 3)
        return rst_fail(CHS_E_ARG, "chs_random_field_tg: bad argument (random_type 1, 2 or 3)");
    // mesh of the turbulence-generator domain alone (CONSPARA; COORDJR.f90:44-55, :166-194)
    const double dx = HX / (double)L, dz = HZ / (double)N;
    std::vector<double> xnd(L + 1), xcc(L), znd(N + 1), zcc(N), ynd(M + 1), ycc(M), dyfi(M), vini(M);
    for (int i = 0; i <= L; ++i) xnd[i] = (double)i * dx;
    for (int i = 0; i < L; ++i) xcc[i] = 0.5 * (xnd[i] + xnd[i + 1]);
    for (int k = 0; k <= N; ++k) znd[k] = (double)k * dz;
    for (int k = 0; k < N; ++k) zcc[k] = 0.5 * (znd[k] + znd[k + 1]);
    for (int j = 0; j <= M; ++j) {
        const double x2 = (double)j / (double)M;
        ynd[j] = std::tanh(STR2 * (x2 - 0.5)) / std::tanh(STR2 * 0.5);
    }
    for (int j = 0; j < M; ++j) {
        ycc[j] = (ynd[j] + ynd[j + 1]) * 0.5;
        dyfi[j] = 1.0 / (ynd[j + 1] - ynd[j]);
        // channel: a = 1, b = 0, PCOE = 3/2
        vini[j] = (ycc[j] < 1.0 && ycc[j] > -1.0) ? (1.0 - (ycc[j] * ycc[j]) / 1.0 / 1.0) * 1.5 : 0.0;
    }
    const size_t hs = (size_t)N * (M + 2) * L;
    auto at = [&](int c, int k, int j, int i) -> double & { return q[c * hs + ((size_t)k * (M + 2) + j) * L + i]; };
    std::fill(q, q + 3 * hs, 0.0);
    std::fill(pr, pr + hs, 0.0);
    // IniField_FLOW_tg: random perturbations (types 1, 2) or the sine field (type 3)
    if (random_type != 3) {
        for (int j = 1; j <= M; ++j) {
            const double vperq = (1.0 - std::fabs(ycc[j - 1])) < 0.25 ? svperg : vperg;
            for (int i = 0; i < L; ++i)
                for (int k = 0; k < N; ++k) {
                    // type 2 reseeds every point with 1973 + I + K + JJ (1-based), so its field is a
                    // function of I + K + JJ; type 1 draws every point independently
                    const uint64_t rd = (uint64_t)(i + 1) + (uint64_t)(k + 1) + (uint64_t)j;
                    const uint64_t key = random_type == 2 ? mix64(seed ^ 0x5bd1e995ull) + 1973ull + rd
                                                          : mix64(seed) ^ ((((uint64_t)j * (uint64_t)N + (uint64_t)k)
                                                                            * (uint64_t)L + (uint64_t)i) + 1ull);
                    double x[3];
                    rvec3(key, x);
                    at(0, k, j, i) = vperq * x[0];
                    at(1, k, j, i) = vperq * x[1];   // RNDI1 = 1 in the channel
                    at(2, k, j, i) = vperq * x[2];   // RCCI1 = 1
                }
            if (j == 1)
                for (int k = 0; k < N; ++k)
                    for (int i = 0; i < L; ++i) at(1, k, j, i) = 0.0;
        }
    } else {
        for (int i = 0; i < L; ++i)
            for (int k = 0; k < N; ++k)
                for (int j = 1; j <= M; ++j) {
                    at(0, k, j, i) = std::sin(xnd[i]) * std::cos(ycc[j - 1]) * std::cos(zcc[k]);
                    at(1, k, j, i) = -std::cos(xcc[i]) * std::sin(ynd[j - 1]) * std::cos(zcc[k]);
                    at(2, k, j, i) = 0.0;
                    pr[((size_t)k * (M + 2) + j) * L + i] = (std::cos(2.0 * xcc[i]) + std::cos(2.0 * ycc[j - 1])) *
                                                            (std::cos(2.0 * zcc[k] + 2.0)) / 16.0;
                }
    }
    // INITIAL_MEAN_TG, then Q = Q - <Q>_xz row by row: u and w averaged at the cell centre
    // ((Q(IP) + Q(I)) / 2, (Q(KP) + Q(K)) / 2), v as stored
    for (int c = 0; c < 3; ++c)
        for (int j = 1; j <= M; ++j) {
            double s = 0.0;
            for (int i = 0; i < L; ++i)
                for (int k = 0; k < N; ++k) {
                    const double ve = c == 0   ? (at(0, k, j, (i + 1) % L) + at(0, k, j, i)) * 0.5
                                      : c == 2 ? (at(2, (k + 1) % N, j, i) + at(2, k, j, i)) * 0.5
                                               : at(1, k, j, i);
                    s = ve + s;
                }
            const double mean = s / (double)(L * N);
            for (int k = 0; k < N; ++k)
                for (int i = 0; i < L; ++i) at(c, k, j, i) = at(c, k, j, i) - mean;
        }
    // u + Vini; v = 0 on the lower wall (row 1); u / bulk velocity; Q * M_inlet / D_inlet
    for (int j = 1; j <= M; ++j)
        for (int i = 0; i < L; ++i)
            for (int k = 0; k < N; ++k) at(0, k, j, i) = at(0, k, j, i) + vini[j - 1];
    for (int k = 0; k < N; ++k)
        for (int i = 0; i < L; ++i) at(1, k, 1, i) = 0.0;
    double um = 0.0;
    for (int j = 1; j <= M; ++j)
        for (int i = 0; i < L; ++i)
            for (int k = 0; k < N; ++k) um = um + at(0, k, j, i) / dyfi[j - 1];
    um = um / (1.0 / dz);
    double area = 0.0;
    for (int j = 0; j < M; ++j)
        for (int k = 0; k < N; ++k) area = area + 1.0 / dyfi[j] / (1.0 / dz);
    um = um / (area * (double)L);
    for (int j = 1; j <= M; ++j)
        for (int k = 0; k < N; ++k)
            for (int i = 0; i < L; ++i) at(0, k, j, i) = at(0, k, j, i) / um;
    for (size_t o = 0; o < 3 * hs; ++o) q[o] = q[o] * m_over_d;
    // INTFC_VARS3 (:93) exchanges halo rows around the periodic rank ring, so when the NS = 0
    // substep starts, the wall halo rows hold the opposite wall's first row: row 0 = row M,
    // row M + 1 = row 1 (that substep's convection reads them before BC_WALL_Q_tg)
    for (int c = 0; c < 3; ++c)
        for (int k = 0; k < N; ++k)
            for (int i = 0; i < L; ++i) {
                at(c, k, 0, i) = at(c, k, M, i);
                at(c, k, M + 1, i) = at(c, k, 1, i);
            }
    return CHS_OK;
}
