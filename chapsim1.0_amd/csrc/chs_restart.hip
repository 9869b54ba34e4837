// Instantaneous-field files of the turbulence-generator domain in the reference's restart
// format (host code): WRT_INSTANT_VARS_TG (processing/WRT_VARS_TG.f90:12-104) writes, per
// variable U, V, W, P, the file <dir>/DNS_perixz_INSTANT_T<PhyTIME as ES15.9>_<var>.D holding
// int32 NCL1, NCL2, NCL3, ITERG, then float64 PhyTIME, REN, DT, then the NCL1 x NCL2 x NCL3
// interior values in Fortran order (MPI native representation, little-endian here);
// ReStart_INSTANT_VARS_TG (processing/RESTART_tg.f90:11-56) with READ_3D_VARS
// (processing/READ_INTP_3D_VARS.f90:20-103) reads them back.
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

}  // namespace

extern "C" {

int chs_restart_write_tg(const char *dir, int L, int M, int N, int iterg, double phytime, double ren, double dt,
                         const double *q, const double *pr)
{
    if (L < 1 || M < 1 || N < 1 || !q || !pr) return rst_fail(CHS_E_ARG, "chs_restart_write_tg: bad argument");
    const size_t hs = (size_t)N * (M + 2) * L;
    std::vector<double> buf((size_t)L * M * N);
    for (int v = 0; v < 4; ++v) {
        const double *f = v < 3 ? q + v * hs : pr;
        // Fortran order (I fastest, then J = 1..M, then K): the halo array's interior rows
        for (int k = 0; k < N; ++k)
            for (int J = 1; J <= M; ++J)
                for (int i = 0; i < L; ++i)
                    buf[((size_t)k * M + (J - 1)) * L + i] = f[((size_t)k * (M + 2) + J) * L + i];
        const std::string fn = rst_name(dir, phytime, kVars[v]);
        FILE *fp = std::fopen(fn.c_str(), "wb");
        if (!fp) return rst_fail(CHS_E_ARG, "chs_restart_write_tg: cannot open " + fn);
        const int32_t hdr[4] = {L, M, N, iterg};
        const double rl[3] = {phytime, ren, dt};
        const bool ok = std::fwrite(hdr, sizeof hdr, 1, fp) == 1 && std::fwrite(rl, sizeof rl, 1, fp) == 1 &&
                        std::fwrite(buf.data(), sizeof(double), buf.size(), fp) == buf.size();
        if (std::fclose(fp) != 0 || !ok) return rst_fail(CHS_E_ARG, "chs_restart_write_tg: write failed: " + fn);
    }
    return CHS_OK;
}

int chs_restart_read_tg(const char *dir, double tfm, int L, int M, int N, int *iterg, double *phytime, double *q,
                        double *pr)
{
    if (L < 1 || M < 1 || N < 1 || !q || !pr) return rst_fail(CHS_E_ARG, "chs_restart_read_tg: bad argument");
    const size_t hs = (size_t)N * (M + 2) * L;
    std::vector<double> buf((size_t)L * M * N);
    for (int v = 0; v < 4; ++v) {
        const std::string fn = rst_name(dir, tfm, kVars[v]);
        FILE *fp = std::fopen(fn.c_str(), "rb");
        if (!fp) return rst_fail(CHS_E_ARG, "chs_restart_read_tg: file " + fn + " does not exist");
        int32_t hdr[4];
        double rl[3];
        const bool ok = std::fread(hdr, sizeof hdr, 1, fp) == 1 && std::fread(rl, sizeof rl, 1, fp) == 1 &&
                        std::fread(buf.data(), sizeof(double), buf.size(), fp) == buf.size();
        std::fclose(fp);
        if (!ok) return rst_fail(CHS_E_ARG, "chs_restart_read_tg: short file " + fn);
        if (hdr[0] != L || hdr[1] != M || hdr[2] != N)
            return rst_fail(CHS_E_ARG, "chs_restart_read_tg: " + fn + " holds " + std::to_string(hdr[0]) + " x " +
                                           std::to_string(hdr[1]) + " x " + std::to_string(hdr[2]));
        if (v == 0) {
            if (iterg) *iterg = hdr[3];
            if (phytime) *phytime = rl[0];
        }
        double *f = v < 3 ? q + v * hs : pr;
        for (int k = 0; k < N; ++k)
            for (int J = 1; J <= M; ++J)
                for (int i = 0; i < L; ++i)
                    f[((size_t)k * (M + 2) + J) * L + i] = buf[((size_t)k * M + (J - 1)) * L + i];
    }
    // walls as FLOWSTART applies them after the read (FLOWSTART.f90:37-40: BC_WALL_Q_tg,
    // BC_WALL_PR_TG): velocity ghost rows and v on J = 1 zero, pressure ghosts copy
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

const char *chs_restart_last_error(void) { return g_rst_err.c_str(); }

}  // extern "C"
