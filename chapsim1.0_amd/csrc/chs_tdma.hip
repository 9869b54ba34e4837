// Batched tridiagonal solvers of the reference (TDMAIJI_CYC, TDMAIJJ_nonCYC, TDMAIJI_nonCYC,
// TDMAIJJ_CYC) on device arrays in the reference's Fortran 2-D layout A, B, C, R(isz, jsz),
// I fastest.  One thread per line; each line follows the reference's recurrence in the same
// operation order with FMA contraction off, so results equal the reference's bit for bit.
// The channel substep does not call these (its line solves are exact-inverse stencils); they
// are the entry points for hosts that solve general systems, e.g. the inflow/outflow x-solves
// (MOMFA_io.f90:147-153) and the triply periodic box Poisson (possion3d_FFT99.f90:355-375).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/chapsim_hip.h"

namespace {

thread_local std::string g_tdma_err;

// work: 3 n doubles per line, element-major ([i][line]) so that consecutive lines coalesce.
__global__ __launch_bounds__(256) void k_tdma_lines(int kind, int isz, int jsz, const double *__restrict__ A,
                                                    const double *__restrict__ B, const double *__restrict__ C,
                                                    double *__restrict__ R, const double *__restrict__ bc,
                                                    double *__restrict__ work)
{
#pragma clang fp contract(off)
    const bool alongI = (kind == 0 || kind == 2), cyc = (kind == 0 || kind == 3);
    const int n = alongI ? isz : jsz, nl = alongI ? jsz : isz;
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nl) return;
    const long es = alongI ? 1 : isz, lstr = alongI ? isz : 1;
    const double *a = A + l * lstr, *b = B + l * lstr, *c = C + l * lstr;
    double *r = R + l * lstr;
    double *H = work, *G1 = work + (size_t)n * nl, *G2 = work + (size_t)2 * n * nl;
    auto E = [&](const double *p, int i) { return p[(long)i * es]; };
    auto W = [&](double *p, int i) -> double & { return p[(size_t)i * nl + l]; };
    if (cyc) {
        // TDMAIJI_CYC.f90:88-157 / TDMAIJJ_CYC.f90:88-157
        W(H, 0) = E(c, 0) / E(b, 0);
        W(G1, 0) = E(r, 0) / E(b, 0);
        W(G2, 0) = -E(a, 0) / E(b, 0);
        for (int i = 1; i <= n - 3; ++i) {
            const double pp = 1.0 / (E(b, i) - W(H, i - 1) * E(a, i));
            const double qq1 = W(G1, i - 1) * E(a, i), qq2 = W(G2, i - 1) * E(a, i);
            W(H, i) = E(c, i) * pp;
            W(G1, i) = (E(r, i) - qq1) * pp;
            W(G2, i) = -qq2 * pp;
        }
        {
            const int i = n - 2;
            const double pp = 1.0 / (E(b, i) - W(H, i - 1) * E(a, i));
            const double qq1 = W(G1, i - 1) * E(a, i), qq2 = W(G2, i - 1) * E(a, i);
            W(G1, i) = (E(r, i) - qq1) * pp;
            W(G2, i) = (-E(c, i) - qq2) * pp;
        }
        for (int i = n - 3; i >= 0; --i) {
            W(G1, i) = -W(H, i) * W(G1, i + 1) + W(G1, i);
            W(G2, i) = -W(H, i) * W(G2, i + 1) + W(G2, i);
        }
        const int e = n - 1;
        const double pp = E(r, e) - E(c, e) * W(G1, 0) - E(a, e) * W(G1, e - 1);
        const double qq1 = E(b, e) + E(c, e) * W(G2, 0) + E(a, e) * W(G2, e - 1);
        const double xe = pp / qq1;
        r[(long)e * es] = xe;
        for (int i = n - 2; i >= 0; --i) r[(long)i * es] = W(G1, i) + W(G2, i) * xe;
    } else {
        // TDMAIJJ_nonCYC.f90:88-123 / TDMAIJI_nonCYC.f90:88-123
        const double bc1 = bc[l], bc2 = bc[nl + l];
        W(H, 0) = E(c, 0) / E(b, 0);
        W(G1, 0) = (E(r, 0) - E(a, 0) * bc1) / E(b, 0);
        for (int i = 1; i <= n - 2; ++i) {
            const double pp = 1.0 / (E(b, i) - W(H, i - 1) * E(a, i));
            W(H, i) = E(c, i) * pp;
            W(G1, i) = (E(r, i) - E(a, i) * W(G1, i - 1)) * pp;
        }
        const int e = n - 1;
        const double pp = 1.0 / (E(b, e) - W(H, e - 1) * E(a, e));
        double x = (E(r, e) - E(c, e) * bc2 - E(a, e) * W(G1, e - 1)) * pp;
        r[(long)e * es] = x;
        for (int i = n - 2; i >= 0; --i) {
            x = W(G1, i) - W(H, i) * x;
            r[(long)i * es] = x;
        }
    }
}

// Work buffer of chs_tdma_lines_noalloc, one per GPU for the life of the process (16 MiB).
// Host code never uses this name as an address: each device's address comes from
// hipGetSymbolAddress (scratch_ptr).
constexpr size_t kScratchDoubles = size_t(1) << 21;
__device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// The current device's address of `scratch`, looked up once per device with it current.
hipError_t scratch_ptr(double **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

}  // namespace

extern "C" {

int chs_tdma_lines(int kind, int isz, int jsz, const double *a, const double *b, const double *c, double *r,
                   const double *bc)
{
    if (kind < 0 || kind > 3) {
        g_tdma_err = "chs_tdma_lines: kind must be 0..3";
        return CHS_E_ARG;
    }
    const bool alongI = (kind == 0 || kind == 2), cyc = (kind == 0 || kind == 3);
    const int n = alongI ? isz : jsz, nl = alongI ? jsz : isz;
    if (isz < 1 || jsz < 1 || n < (cyc ? 3 : 2) || !a || !b || !c || !r || (!cyc && !bc)) {
        g_tdma_err = "chs_tdma_lines: bad size or null array (lines need 3 points cyclic, 2 otherwise)";
        return CHS_E_ARG;
    }
    double *work = nullptr;
    hipError_t e = hipMalloc(&work, sizeof(double) * 3 * (size_t)n * nl);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_tdma_lines, dim3((nl + 255) / 256), dim3(256), 0, 0, kind, isz, jsz, a, b, c, r,
                           cyc ? nullptr : bc, work);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(0);
        (void)hipFree(work);
    }
    if (e != hipSuccess) {
        g_tdma_err = std::string("chs_tdma_lines: ") + hipGetErrorString(e);
        return CHS_E_HIP;
    }
    return CHS_OK;
}

int chs_tdma_lines_noalloc(int kind, int isz, int jsz, const double *a, const double *b, const double *c, double *r,
                           const double *bc)
{
    if (kind < 0 || kind > 3) {
        g_tdma_err = "chs_tdma_lines_noalloc: kind must be 0..3";
        return CHS_E_ARG;
    }
    const bool alongI = (kind == 0 || kind == 2), cyc = (kind == 0 || kind == 3);
    const int n = alongI ? isz : jsz, nl = alongI ? jsz : isz;
    if (isz < 1 || jsz < 1 || n < (cyc ? 3 : 2) || !a || !b || !c || !r || (!cyc && !bc)) {
        g_tdma_err = "chs_tdma_lines_noalloc: bad size or null array (lines need 3 points cyclic, 2 otherwise)";
        return CHS_E_ARG;
    }
    if (3 * (size_t)n * (size_t)nl > kScratchDoubles) {
        g_tdma_err = "chs_tdma_lines_noalloc: 3 * points per line * lines exceeds the " +
                     std::to_string(kScratchDoubles) + "-double work buffer";
        return CHS_E_ARG;
    }
    // every work entry a line reads was written earlier by the same line: nothing to clear
    double *work = nullptr;
    hipError_t e = scratch_ptr(&work);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_tdma_lines, dim3((nl + 255) / 256), dim3(256), 0, 0, kind, isz, jsz, a, b, c, r,
                           cyc ? nullptr : bc, work);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(0);
    }
    if (e != hipSuccess) {
        g_tdma_err = std::string("chs_tdma_lines_noalloc: ") + hipGetErrorString(e);
        return CHS_E_HIP;
    }
    return CHS_OK;
}

const char *chs_tdma_last_error(void) { return g_tdma_err.c_str(); }

}  // extern "C"
