// Register-resident complex FFTs of power-of-two length for gfx950.
//
// One line of N points is owned by TL = N / E threads of one wavefront; each thread holds E
// points in registers.  A pass of radix R (8 or 16, or 4 as the last pass) runs E / R
// butterflies per thread in registers; between passes the line goes once through LDS
// (padded against bank conflicts), with only a wavefront-level ordering barrier, because a
// wavefront never shares a line with another.  Stockham autosort order: pass p with
// Ns = R_0 ... R_{p-1} maps butterfly j (k = j mod Ns) inputs x[j + r N / R] (times the twiddle
// w^(r k), w = exp(-+2 pi i / (Ns R))) to outputs y[(j / Ns) Ns R + k + r Ns]; the last pass
// writes natural order.  Twiddles of pass p are a table [r - 1][k] (forward sign) so that
// consecutive threads read consecutive entries.
#pragma once
#include <cmath>
#include <vector>

#include "chs_fft.h"

namespace chs {

constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n / 2); }

#ifdef CHS_HOST_EMU
void wave_lds_sync();   // host emulation (lanes as threads): a barrier of the line's lanes
#else
// Orders the wavefront's LDS accesses across lanes: LDS instructions of one wavefront
// execute in issue order, so only compiler reordering has to be prevented.
__device__ __forceinline__ void wave_lds_sync() { __builtin_amdgcn_wave_barrier(); }
#endif

template <typename T> __device__ __forceinline__ cpx<T> crot(cpx<T> a, T c, T s)
{
    return {a.x * c - a.y * s, a.x * s + a.y * c};
}

template <typename T> __device__ __forceinline__ void dft4(cpx<T> &x0, cpx<T> &x1, cpx<T> &x2, cpx<T> &x3, int sign)
{
    const cpx<T> a0 = cadd(x0, x2), a1 = csub(x0, x2);
    const cpx<T> b0 = cadd(x1, x3), b1 = cmuli(csub(x1, x3), sign);
    x0 = cadd(a0, b0);
    x2 = csub(a0, b0);
    x1 = cadd(a1, b1);
    x3 = csub(a1, b1);
}

// In-register DFT of length 8 or 16 (natural order in and out), as 4 x 2 and 4 x 4:
// X[k1 + 4 k2] = sum_n2 w_R^(n2 k1) w_(R/4)^(n2 k2) sum_n1 x[(R/4) n1 + n2] w_4^(n1 k1).
template <typename T, int R> __device__ __forceinline__ void rdft(cpx<T> *v, int sign)
{
    static_assert(R == 4 || R == 8 || R == 16, "radix");
    if constexpr (R == 4) {
        dft4(v[0], v[1], v[2], v[3], sign);
    } else if constexpr (R == 8) {
        constexpr T h = T(0.70710678118654752440084436210485);
        const T s = T(sign);
        dft4(v[0], v[2], v[4], v[6], sign);   // n2 = 0 -> k1 at v[2 k1]
        dft4(v[1], v[3], v[5], v[7], sign);   // n2 = 1 -> k1 at v[2 k1 + 1]
        v[3] = crot(v[3], h, s * h);          // w8^1
        v[5] = cmuli(v[5], sign);             // w8^2
        v[7] = crot(v[7], -h, s * h);         // w8^3
        cpx<T> o[8];
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            o[k1] = cadd(v[2 * k1], v[2 * k1 + 1]);
            o[k1 + 4] = csub(v[2 * k1], v[2 * k1 + 1]);
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = o[q];
    } else {
        constexpr T c1 = T(0.92387953251128675612818318939679), s1 = T(0.38268343236508977172845998403040);
        constexpr T h = T(0.70710678118654752440084436210485);
        const T s = T(sign);
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) dft4(v[n2], v[4 + n2], v[8 + n2], v[12 + n2], sign);   // k1 at v[4 k1 + n2]
        // twiddles w16^(n2 k1) on v[4 k1 + n2]
        v[5] = crot(v[5], c1, s * s1);      // 1
        v[6] = crot(v[6], h, s * h);        // 2
        v[7] = crot(v[7], s1, s * c1);      // 3
        v[9] = crot(v[9], h, s * h);        // 2
        v[10] = cmuli(v[10], sign);         // 4
        v[11] = crot(v[11], -h, s * h);     // 6
        v[13] = crot(v[13], s1, s * c1);    // 3
        v[14] = crot(v[14], -h, s * h);     // 6
        v[15] = crot(v[15], -c1, -s * s1);  // 9
        cpx<T> o[16];
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            cpx<T> a = v[4 * k1], b = v[4 * k1 + 1], c = v[4 * k1 + 2], d = v[4 * k1 + 3];
            dft4(a, b, c, d, sign);
            o[k1] = a;
            o[k1 + 4] = b;
            o[k1 + 8] = c;
            o[k1 + 12] = d;
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = o[q];
    }
}

// Padded LDS index of line element o (one slot of padding every 2^PS elements).
template <int PS> __device__ __forceinline__ int rpad(int o) { return o + (o >> PS); }
template <int N, int PS> constexpr int rline_stride() { return N + (N >> PS) + 1; }

template <int A, int... B> struct first_of { static constexpr int value = A; };

// One pass and the rest, recursively.  v holds this pass's inputs: v[b R + r] =
// x[t + b TL + r N / R].  Off = offset of this pass's twiddles in the table.  The last pass
// hands each output (natural index, value) to out and leaves it in v.  lds is the line's
// buffer: cpx<T>[LS], or with SPLIT T[LS], through which the real parts go first and then the
// imaginary parts (half the LDS of a line, two more wavefront barriers per exchange).
template <bool SPLIT, typename T, int N, int E, int PS, int Ns, int Off, int R, int... Rest, typename Out>
__device__ __forceinline__ void rfft_pass(cpx<T> (&v)[E], void *lds, int t, const cpx<T> *__restrict__ tw,
                                          int sign, Out &&out)
{
    constexpr int TL = N / E, NB = E / R;
    static_assert(E % R == 0 && N % (Ns * R) == 0, "plan");
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int j = t + b * TL;
        const int k = j % Ns;
        cpx<T> *x = &v[b * R];
        if constexpr (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                cpx<T> w = tw[Off + (r - 1) * Ns + k];
                if (sign > 0) w.y = -w.y;
                x[r] = cmul(x[r], w);
            }
        }
        rdft<T, R>(x, sign);
    }
    if constexpr (sizeof...(Rest) == 0) {
        static_assert(Ns * R == N, "plan");
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int r = 0; r < R; ++r) out(t + b * TL + r * Ns, v[b * R + r]);
    } else {
        constexpr int R2 = first_of<Rest...>::value;
        constexpr int NB2 = E / R2, NR2 = N / R2;
        auto slot = [&](int b, int r) { return ((t + b * TL) / Ns) * Ns * R + (t + b * TL) % Ns + r * Ns; };
        if constexpr (!SPLIT) {
            cpx<T> *l = static_cast<cpx<T> *>(lds);
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int r = 0; r < R; ++r) l[rpad<PS>(slot(b, r))] = v[b * R + r];
            wave_lds_sync();
#pragma unroll
            for (int b = 0; b < NB2; ++b)
#pragma unroll
                for (int r = 0; r < R2; ++r) v[b * R2 + r] = l[rpad<PS>(t + b * TL + r * NR2)];
            wave_lds_sync();
        } else {
            T *l = static_cast<T *>(lds);
            T xin[E];
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int r = 0; r < R; ++r) l[rpad<PS>(slot(b, r))] = v[b * R + r].x;
            wave_lds_sync();
#pragma unroll
            for (int b = 0; b < NB2; ++b)
#pragma unroll
                for (int r = 0; r < R2; ++r) xin[b * R2 + r] = l[rpad<PS>(t + b * TL + r * NR2)];
            wave_lds_sync();
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int r = 0; r < R; ++r) l[rpad<PS>(slot(b, r))] = v[b * R + r].y;
            wave_lds_sync();
#pragma unroll
            for (int b = 0; b < NB2; ++b)
#pragma unroll
                for (int r = 0; r < R2; ++r) v[b * R2 + r] = {xin[b * R2 + r], l[rpad<PS>(t + b * TL + r * NR2)]};
            wave_lds_sync();
        }
        rfft_pass<SPLIT, T, N, E, PS, Ns * R, (Ns > 1 ? Off + Ns * (R - 1) : Off), Rest...>(v, lds, t, tw, sign, out);
    }
}

// Plans: N = product of the radices; E points per thread (E divisible by every radix).
template <int N> struct RPlan;
template <> struct RPlan<64> { static constexpr int E = 8, PS = 3; };
template <> struct RPlan<128> { static constexpr int E = 16, PS = 4; };
template <> struct RPlan<256> { static constexpr int E = 16, PS = 4; };
template <> struct RPlan<512> { static constexpr int E = 8, PS = 3; };
template <> struct RPlan<1024> { static constexpr int E = 16, PS = 4; };

// Full transform of one line: v holds the first pass's inputs (v[b R0 + r] =
// x[t + b TL + r N / R0], TL = N / E); out(index, value) receives the natural-order result,
// which also stays in v (v[e] is element rout_index<N>(t, e)).  lds: cpx<T>[LS], or T[LS] with
// SPLIT.
template <typename T, int N, bool SPLIT = false, typename Out>
__device__ __forceinline__ void rfft_line(cpx<T> (&v)[RPlan<N>::E], void *lds, int t, const cpx<T> *__restrict__ tw,
                                          int sign, Out &&out)
{
    constexpr int E = RPlan<N>::E, PS = RPlan<N>::PS;
    if constexpr (N == 64) rfft_pass<SPLIT, T, N, E, PS, 1, 0, 8, 8>(v, lds, t, tw, sign, out);
    else if constexpr (N == 128) rfft_pass<SPLIT, T, N, E, PS, 1, 0, 16, 8>(v, lds, t, tw, sign, out);
    else if constexpr (N == 256) rfft_pass<SPLIT, T, N, E, PS, 1, 0, 16, 16>(v, lds, t, tw, sign, out);
    else if constexpr (N == 512) rfft_pass<SPLIT, T, N, E, PS, 1, 0, 8, 8, 8>(v, lds, t, tw, sign, out);
    else rfft_pass<SPLIT, T, N, E, PS, 1, 0, 16, 16, 4>(v, lds, t, tw, sign, out);
}

// Last radix of the plan for N, and the natural index of output slot e of lane t.
template <int N> constexpr int rlast() { return N == 128 ? 8 : (N == 1024 ? 4 : (N == 512 ? 8 : (N == 64 ? 8 : 16))); }
template <int N> __device__ __forceinline__ int rout_index(int t, int e)
{
    constexpr int R = rlast<N>(), TL = N / RPlan<N>::E, Ns = N / R;
    return t + (e / R) * TL + (e % R) * Ns;
}

// First radix of the plan (pass-0 input stride is N / rfirst).
template <int N> constexpr int rfirst() { return (N == 64 || N == 512) ? 8 : 16; }

// Radices of the plan for n (host side; empty if n has no register plan).
inline int rplan_radices(int n, int *r)
{
    switch (n) {
    case 64: r[0] = 8; r[1] = 8; return 2;
    case 128: r[0] = 16; r[1] = 8; return 2;
    case 256: r[0] = 16; r[1] = 16; return 2;
    case 512: r[0] = 8; r[1] = 8; r[2] = 8; return 3;
    case 1024: r[0] = 16; r[1] = 16; r[2] = 4; return 3;
    default: return 0;
    }
}

// Points per thread and lines per wavefront of the plan for n (host side).
inline int rplan_e(int n) { return (n == 64 || n == 512) ? 8 : 16; }
inline int rplan_lpw(int n) { return 64 / (n / rplan_e(n)); }

// Twiddle tables of passes 1.. of the plan for n, concatenated: pass p (Ns = R_0 ... R_{p-1},
// radix R) holds exp(-2 pi i r k / (Ns R)) at [(r - 1) Ns + k], r = 1..R-1, k = 0..Ns-1.
// Computed in long double.
template <typename T> std::vector<cpx<T>> rplan_twiddles(int n)
{
    int rad[8];
    const int np = rplan_radices(n, rad);
    std::vector<cpx<T>> tw;
    const long double pi = 3.141592653589793238462643383279502884L;
    int Ns = np > 0 ? rad[0] : 1;
    for (int p = 1; p < np; ++p) {
        const int R = rad[p];
        for (int r = 1; r < R; ++r)
            for (int k = 0; k < Ns; ++k) {
                const long double a = -2.0L * pi * (long double)((long long)r * k) / (long double)(Ns * R);
                tw.push_back({(T)std::cos(a), (T)std::sin(a)});
            }
        Ns *= R;
    }
    return tw;
}

}  // namespace chs
