// Host emulation of the register FFT of chapsim1.0_amd/csrc/chs_fftr.h: the lanes that own one
// line run as threads and meet at a barrier wherever the device code orders its LDS accesses
// (wave_lds_sync).  Each planned length is checked in both directions against a long-double
// DFT; prints one line per case and exits non-zero on a failure.
#define CHS_HOST_EMU 1
#define __device__
#define __forceinline__ inline
#define __restrict__

#include <barrier>
#include <cmath>
#include <cstdio>
#include <random>
#include <thread>
#include <vector>

static thread_local std::barrier<> *g_bar = nullptr;
namespace chs {
void wave_lds_sync() { g_bar->arrive_and_wait(); }
}  // namespace chs

#include "../../chapsim1.0_amd/csrc/chs_fftr.h"

template <typename T, int N, bool SPLIT> double run(int sign, unsigned seed)
{
    using C = chs::cpx<T>;
    constexpr int E = chs::RPlan<N>::E, PS = chs::RPlan<N>::PS, TL = N / E, R0 = chs::rfirst<N>();
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    std::vector<C> x(N), out(N, C{T(0), T(0)});
    for (auto &e : x) e = {(T)U(rng), (T)U(rng)};
    const std::vector<C> tw = chs::rplan_twiddles<T>(N);
    std::vector<C> lds(chs::rline_stride<N, PS>());   // SPLIT uses the first half as T[LS]
    std::barrier<> bar(TL);
    std::vector<std::thread> th;
    for (int t = 0; t < TL; ++t)
        th.emplace_back([&, t] {
            g_bar = &bar;
            C v[E];
            for (int b = 0; b < E / R0; ++b)
                for (int r = 0; r < R0; ++r) v[b * R0 + r] = x[t + b * TL + r * (N / R0)];
            chs::rfft_line<T, N, SPLIT>(v, lds.data(), t, tw.data(), sign, [&](int idx, C val) { out[idx] = val; });
            // the outputs also stay in v, slot e at its natural index rout_index
            for (int e = 0; e < E; ++e) {
                const C w = out[chs::rout_index<N>(t, e)];
                if (w.x != v[e].x || w.y != v[e].y) out[chs::rout_index<N>(t, e)] = {T(1e30), T(1e30)};
            }
        });
    for (auto &h : th) h.join();
    const long double pi = 3.141592653589793238462643383279502884L;
    long double err = 0, mag = 0;
    for (int k = 0; k < N; ++k) {
        long double re = 0, im = 0;
        for (int n = 0; n < N; ++n) {
            const long double a = sign * 2.0L * pi * (long double)((long long)n * k % N) / N;
            re += x[n].x * std::cos(a) - x[n].y * std::sin(a);
            im += x[n].x * std::sin(a) + x[n].y * std::cos(a);
        }
        err = std::fmax(err, std::fabs(re - out[k].x));
        err = std::fmax(err, std::fabs(im - out[k].y));
        mag = std::fmax(mag, std::fmax(std::fabs(re), std::fabs(im)));
    }
    return (double)(err / mag);
}

template <int N> int check()
{
    int bad = 0;
    for (int sign : {-1, 1}) {
        for (int split = 0; split < 2; ++split) {
            const double e64 = split ? run<double, N, true>(sign, 7 + N) : run<double, N, false>(sign, 7 + N);
            const double e32 = split ? run<float, N, true>(sign, 11 + N) : run<float, N, false>(sign, 11 + N);
            const bool ok = e64 < 1e-14 && e32 < 2e-6;
            std::printf("N=%d sign=%+d split=%d fp64 %.3e fp32 %.3e %s\n", N, sign, split, e64, e32, ok ? "ok" : "FAIL");
            bad += !ok;
        }
    }
    return bad;
}

int main()
{
    int bad = check<64>() + check<128>() + check<256>() + check<512>() + check<1024>();
    return bad ? 1 : 0;
}
