// Thermophysical property tables for the heated channel (table mode, iThermoProperty =
// search_table): host restatement of the reference's table set-up and a GPU kernel for the
// per-cell update THERM_PROP_UPDATE_FROM_DH (THERM_PROP_UPDATE.f90:14-67).
//
// Bit-exactness: every lookup is a bisection (integer) followed by a Horner evaluation in
// FP64 with the reference's operation order and no fused multiply-add (contraction is
// switched off below), so results equal the reference build's bit for bit; the table set-up
// is restated in the same operation order for the same reason.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
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

// spline_interpolation_* (thermo_properties_table_searching.f90:678-723 and siblings):
// range check (tolerance 1e-8), bisection I = 1, J = N+1, then
// Y(I) + DX*(B(I) + DX*(C(I) + DX*D(I))).
__device__ double spl_eval(const SplineDev &s, int n, double xv, int *err, int bit)
{
#pragma clang fp contract(off)
    if ((xv - s.x[s.imin - 1]) < -1.0e-8) atomicOr(err, bit);
    if ((xv - s.x[s.imax - 1]) > 1.0e-8) atomicOr(err, bit << 8);
    int I = 1, J = n + 1;
    while (J > I + 1) {
        const int K = (I + J) / 2;
        if (xv < s.x[K - 1]) J = K;
        else I = K;
    }
    const double dx = xv - s.x[I - 1];
    return s.y[I - 1] + dx * (s.b[I - 1] + dx * (s.c[I - 1] + dx * s.d[I - 1]));
}

struct PropDev {
    int n;
    SplineDev dhh, ht, hd, hm, hk, hcp, hb;
};

// THERM_PROP_UPDATE_FROM_DH, table branch (THERM_PROP_UPDATE.f90:45-52), one thread per
// value.  D is taken from the H->M spline, M from H->K and K from H->D, as the reference
// does (lines 49-51).  out = [H | T | D | M | K | Cp | B], each n values.
__global__ __launch_bounds__(256) void k_prop_from_dh(PropDev p, long n, const double *__restrict__ dh,
                                                      double *__restrict__ out, int *__restrict__ err)
{
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    const double h = spl_eval(p.dhh, p.n, dh[t], err, 1);
    out[t] = h;
    out[n + t] = spl_eval(p.ht, p.n, h, err, 2);
    out[2 * n + t] = spl_eval(p.hm, p.n, h, err, 4);
    out[3 * n + t] = spl_eval(p.hk, p.n, h, err, 8);
    out[4 * n + t] = spl_eval(p.hd, p.n, h, err, 16);
    out[5 * n + t] = spl_eval(p.hcp, p.n, h, err, 32);
    out[6 * n + t] = spl_eval(p.hb, p.n, h, err, 64);
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

}  // namespace

struct chs_prop {
    int n = 0;
    double *dev = nullptr;   // all table arrays, packed
    int *err = nullptr;
    PropDev p;
    hipStream_t stream = nullptr;
    std::string errmsg;
};

extern "C" {

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
    *out = p;
    return CHS_OK;
}

void chs_prop_destroy(chs_prop *p)
{
    if (!p) return;
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    if (p->dev) (void)hipFree(p->dev);
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
