// Batched complex FFTs staged in LDS (Stockham autosort, mixed radix 2/3/4/5) for gfx950.
// Used by the pressure Poisson solve: real x-transforms (two real lines packed into one
// complex line) and complex z-transforms.  Twiddles come from a table w[m] = exp(-2 pi i
// m / n) held in LDS; the inverse uses its conjugate.  The transforms are unnormalised.
#pragma once
#ifndef CHS_HOST_EMU
#include <hip/hip_runtime.h>
#endif

namespace chs {

template <typename T> struct cpx { T x, y; };

template <typename T> __device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T> __device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T> __device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
// multiply by (sign * i)
template <typename T> __device__ __forceinline__ cpx<T> cmuli(cpx<T> a, int sign)
{
    return sign < 0 ? cpx<T>{a.y, -a.x} : cpx<T>{-a.y, a.x};
}

// In-register DFT of length R (2, 3, 4, 5) with exponent sign `sign`.
template <typename T, int R> __device__ __forceinline__ void bfly(cpx<T> *v, int sign)
{
    if constexpr (R == 2) {
        cpx<T> a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 4) {
        cpx<T> a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
        cpx<T> b0 = cadd(v[1], v[3]), b1 = cmuli(csub(v[1], v[3]), sign);
        v[0] = cadd(a0, b0);
        v[2] = csub(a0, b0);
        v[1] = cadd(a1, b1);
        v[3] = csub(a1, b1);
    } else if constexpr (R == 3) {
        const T c1 = T(-0.5), s1 = T(sign) * T(0.86602540378443864676372317075294);
        cpx<T> t = cadd(v[1], v[2]);
        cpx<T> d = csub(v[1], v[2]);
        cpx<T> m = {v[0].x + c1 * t.x, v[0].y + c1 * t.y};
        cpx<T> r = {-s1 * d.y, s1 * d.x};
        v[0] = cadd(v[0], t);
        v[1] = cadd(m, r);
        v[2] = csub(m, r);
    } else {   // R == 5
        const T c1 = T(0.30901699437494742410229341718282), c2 = T(-0.80901699437494742410229341718282);
        const T s1 = T(sign) * T(0.95105651629515357211643933337938);
        const T s2 = T(sign) * T(0.58778525229247312916870595463907);
        cpx<T> t1 = cadd(v[1], v[4]), d1 = csub(v[1], v[4]);
        cpx<T> t2 = cadd(v[2], v[3]), d2 = csub(v[2], v[3]);
        cpx<T> m1 = {v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y};
        cpx<T> m2 = {v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y};
        // i * (s1 d1 + s2 d2) and i * (s2 d1 - s1 d2)
        cpx<T> r1 = {-(s1 * d1.y + s2 * d2.y), s1 * d1.x + s2 * d2.x};
        cpx<T> r2 = {-(s2 * d1.y - s1 * d2.y), s2 * d1.x - s1 * d2.x};
        v[0] = {v[0].x + t1.x + t2.x, v[0].y + t1.y + t2.y};
        v[1] = cadd(m1, r1);
        v[4] = csub(m1, r1);
        v[2] = cadd(m2, r2);
        v[3] = csub(m2, r2);
    }
}

struct FftPlan {
    int n;
    int npass;
    int radix[16];
};

#ifndef CHS_HOST_EMU
// Global -> LDS staging of n items by a block: ld(t) reads item t, st(t, v) stores it.  The U
// reads of one thread are all issued before the first store, so U memory latencies overlap
// instead of following one another.
template <int U, typename V, typename Ld, typename St>
__device__ __forceinline__ void stage_items(int n, Ld ld, St st)
{
    const int nthr = blockDim.x * blockDim.y;
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    for (int t0 = tid; t0 < n; t0 += U * nthr) {
        V v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int t = t0 + u * nthr;
            if (t < n) v[u] = ld(t);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int t = t0 + u * nthr;
            if (t < n) st(t, v[u]);
        }
    }
}

#endif

// Largest number of line points per thread of a block (nlines * n <= kMaxPts * nthr); a
// radix-R pass then holds ceil(kMaxPts / R) butterflies of one thread in registers.
constexpr int kMaxPts = 8;

#ifndef CHS_HOST_EMU
template <typename T, int R>
__device__ __forceinline__ void fft_pass_inplace(cpx<T> *a, int ls, int nlines, int n, int Ns,
                                                 const cpx<T> *tw, int sign, int tid, int nthr)
{
    const int nb = n / R;
    const int tstride = n / (Ns * R);
    const int total = nb * nlines;
    constexpr int kB = (kMaxPts + R - 1) / R;
    cpx<T> v[kB][R];
#pragma unroll
    for (int b = 0; b < kB; ++b) {
        const int t = tid + b * nthr;
        if (t < total) {
            const int line = t / nb;
            const int j = t - line * nb;
            const int k = j % Ns;
            const cpx<T> *src = a + line * ls;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                cpx<T> x = src[j + r * nb];
                if (r > 0) {
                    cpx<T> w = tw[r * k * tstride];
                    if (sign > 0) w.y = -w.y;
                    x = cmul(x, w);
                }
                v[b][r] = x;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < kB; ++b) {
        const int t = tid + b * nthr;
        if (t < total) {
            const int line = t / nb;
            const int j = t - line * nb;
            const int k = j % Ns;
            bfly<T, R>(v[b], sign);
            cpx<T> *dst = a + line * ls;
            const int idxD = (j / Ns) * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) dst[idxD + r * Ns] = v[b][r];
        }
    }
    __syncthreads();
}

// In-place Stockham transform of `nlines` lines of length plan.n stored at
// a[line * ls + idx] in LDS; tw holds exp(-2 pi i m / n), m = 0..n-1 (LDS).  Each pass reads
// all butterfly inputs into registers, synchronises, then writes.  Requires
// nlines * n <= kMaxPts * nthr.  All threads of the block must call.
template <typename T>
__device__ void fft_lds_inplace(cpx<T> *a, int ls, int nlines, const FftPlan &plan,
                                const cpx<T> *tw, int sign, int tid, int nthr)
{
    int Ns = 1;
    for (int p = 0; p < plan.npass; ++p) {
        const int R = plan.radix[p];
        switch (R) {
        case 4: fft_pass_inplace<T, 4>(a, ls, nlines, plan.n, Ns, tw, sign, tid, nthr); break;
        case 2: fft_pass_inplace<T, 2>(a, ls, nlines, plan.n, Ns, tw, sign, tid, nthr); break;
        case 3: fft_pass_inplace<T, 3>(a, ls, nlines, plan.n, Ns, tw, sign, tid, nthr); break;
        default: fft_pass_inplace<T, 5>(a, ls, nlines, plan.n, Ns, tw, sign, tid, nthr); break;
        }
        Ns *= R;
    }
}

#endif

}  // namespace chs
