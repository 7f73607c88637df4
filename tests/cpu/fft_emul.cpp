// CPU emulation of the device line FFT (mpc_amd/csrc/fft_core.h): runs the
// per-thread stage code in a loop over threads between "barriers" and
// compares every plan size against a brute-force DFT in long double.
// Prints the max relative error per (N, precision, direction); exit 1 on failure.
#include "../../mpc_amd/csrc/fft_core.h"
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

using namespace mpc;

template <class T, int N, int E, int DIR, int PAD, int Ns, int R, int... Rest>
struct HostRun {
    static void run(std::vector<cx<T>> &V, std::vector<cx<T>> &lds, const cx<T> *tw) {
        constexpr int TPL = N / E;
        constexpr bool last = sizeof...(Rest) == 0;
        for (int tid = 0; tid < TPL; ++tid) {
            cx<T>(&v)[E] = *reinterpret_cast<cx<T>(*)[E]>(&V[(size_t)tid * E]);
            fft_stage<T, N, E, R, Ns, DIR, last, PAD>(v, tid, lds.data(), 1, tw);
        }
        if constexpr (!last) {
            for (int tid = 0; tid < TPL; ++tid) {
                cx<T>(&v)[E] = *reinterpret_cast<cx<T>(*)[E]>(&V[(size_t)tid * E]);
                fft_gather<T, N, E, PAD>(v, tid, lds.data(), 1);
            }
            HostRun<T, N, E, DIR, PAD, Ns * R, Rest...>::run(V, lds, tw);
        }
    }
};
template <class T, int N, int E, int DIR, int PAD, class Rad> struct Unpack;
template <class T, int N, int E, int DIR, int PAD, int... Rs> struct Unpack<T, N, E, DIR, PAD, Radices<Rs...>> {
    static void run(std::vector<cx<T>> &V, std::vector<cx<T>> &lds, const cx<T> *tw) {
        HostRun<T, N, E, DIR, PAD, 1, Rs...>::run(V, lds, tw);
    }
};

/* the same transform with each thread's twiddle powers loaded once into a
 * TwChain (the fused t-line kernel's form) */
template <class T, int N, int E, int DIR, int PAD, int Ns, int R, int... Rest>
struct HostRunC {
    using CH = TwChain<T, N, E, Ns, R, Rest...>;
    static void run(std::vector<cx<T>> &V, std::vector<cx<T>> &lds, const std::vector<const CH *> &ch) {
        constexpr int TPL = N / E;
        constexpr bool last = sizeof...(Rest) == 0;
        for (int tid = 0; tid < TPL; ++tid) {
            cx<T>(&v)[E] = *reinterpret_cast<cx<T>(*)[E]>(&V[(size_t)tid * E]);
            fft_stage_t<T, N, E, R, Ns, DIR, last, PAD>(v, tid, lds.data(), 1,
                                                         TwFromPowers<T, R, E / R, DIR>{ch[tid]->p});
        }
        if constexpr (!last) {
            for (int tid = 0; tid < TPL; ++tid) {
                cx<T>(&v)[E] = *reinterpret_cast<cx<T>(*)[E]>(&V[(size_t)tid * E]);
                fft_gather<T, N, E, PAD>(v, tid, lds.data(), 1);
            }
            std::vector<const TwChain<T, N, E, Ns * R, Rest...> *> nx;
            for (int tid = 0; tid < TPL; ++tid) nx.push_back(&ch[tid]->next);
            HostRunC<T, N, E, DIR, PAD, Ns * R, Rest...>::run(V, lds, nx);
        }
    }
};
template <class T, int N, int E, int DIR, int PAD, class Rad> struct UnpackC;
template <class T, int N, int E, int DIR, int PAD, int... Rs> struct UnpackC<T, N, E, DIR, PAD, Radices<Rs...>> {
    static void run(std::vector<cx<T>> &V, std::vector<cx<T>> &lds, const cx<T> *tw) {
        constexpr int TPL = N / E;
        std::vector<TwChain<T, N, E, 1, Rs...>> chains(TPL);
        std::vector<const TwChain<T, N, E, 1, Rs...> *> ptr;
        for (int tid = 0; tid < TPL; ++tid) { chains[tid].load(tw, tid); ptr.push_back(&chains[tid]); }
        HostRunC<T, N, E, DIR, PAD, 1, Rs...>::run(V, lds, ptr);
    }
};

template <class T, int N, int DIR, class PL = Plan<N>, int PAD = 0, bool CACHED = false> double check(unsigned seed) {
    constexpr int E = PL::E, TPL = N / E;
    std::vector<cx<T>> tw(N), lds(padded_elems<PAD>(N)), V((size_t)TPL * E);
    for (int k = 0; k < N; ++k) {
        long double a = -2.0L * 3.14159265358979323846264338327950288L * k / N;
        tw[k] = mk((T)cosl(a), (T)sinl(a));
    }
    std::vector<long double> xr(N), xi(N);
    srand(seed);
    for (int i = 0; i < N; ++i) { xr[i] = (T)(rand() / (double)RAND_MAX - 0.5); xi[i] = (T)(rand() / (double)RAND_MAX - 0.5); }
    for (int tid = 0; tid < TPL; ++tid)
        for (int m = 0; m < E; ++m) V[(size_t)tid * E + m] = mk((T)xr[tid + TPL * m], (T)xi[tid + TPL * m]);
    if (CACHED) UnpackC<T, N, E, DIR, PAD, typename PL::R>::run(V, lds, tw.data());
    else Unpack<T, N, E, DIR, PAD, typename PL::R>::run(V, lds, tw.data());
    double num = 0, den = 0;
    for (int k = 0; k < N; ++k) {
        long double sr = 0, si = 0;
        for (int j = 0; j < N; ++j) {
            long double a = DIR * 2.0L * 3.14159265358979323846264338327950288L * (long double)(((long)j * k) % N) / N;
            sr += xr[j] * cosl(a) - xi[j] * sinl(a);
            si += xr[j] * sinl(a) + xi[j] * cosl(a);
        }
        const int tid = k % TPL, m = k / TPL;
        const cx<T> g = V[(size_t)tid * E + m];
        num += (double)((g.re - sr) * (g.re - sr) + (g.im - si) * (g.im - si));
        den += (double)(sr * sr + si * si);
    }
    return std::sqrt(num / den);
}

template <int N> int run_n() {
    int bad = 0;
    double e1 = check<double, N, -1>(N), e2 = check<double, N, +1>(N + 1);
    double e3 = check<float, N, -1>(N + 2), e4 = check<float, N, +1>(N + 3);
    printf("N=%5d f64 fwd %.3e inv %.3e  f32 fwd %.3e inv %.3e\n", N, e1, e2, e3, e4);
    if (e1 > 1e-14 || e2 > 1e-14 || e3 > 2e-6 || e4 > 2e-6) bad = 1;
    return bad;
}

template <int N> int run_t() {
    /* t-line plans with the padded LDS layout used by the fused kernel */
    double e1 = check<double, N, -1, TPlan<N>, 8>(N), e2 = check<double, N, +1, TPlan<N>, 8>(N + 1);
    double e3 = check<float, N, -1, TPlan<N>, 8>(N + 2), e4 = check<float, N, +1, TPlan<N>, 8>(N + 3);
    printf("t-plan N=%5d f64 fwd %.3e inv %.3e  f32 fwd %.3e inv %.3e\n", N, e1, e2, e3, e4);
    int bad = (e1 > 1e-14 || e2 > 1e-14 || e3 > 2e-6 || e4 > 2e-6) ? 1 : 0;
    /* the swizzled layout the t-line kernel uses */
    double s1 = check<double, N, -1, TPlan<N>, -3>(N), s2 = check<float, N, +1, TPlan<N>, -3>(N + 3);
    printf("swizzled N=%5d f64 fwd %.3e  f32 inv %.3e\n", N, s1, s2);
    if (s1 != e1 || s2 != e4) { printf("swizzled layout changes the result at N=%d\n", N); bad = 1; }
    /* cached twiddle powers, both t plans */
    double c1 = check<double, N, -1, TPlan<N>, 8, true>(N), c2 = check<double, N, +1, TPlan<N>, 8, true>(N + 1);
    double c3 = check<float, N, -1, Plan<N>, -4, true>(N + 2), c4 = check<float, N, +1, Plan<N>, -4, true>(N + 3);
    printf("cached N=%5d f64 fwd %.3e inv %.3e  f32 (E=16) fwd %.3e inv %.3e\n", N, c1, c2, c3, c4);
    if (c1 > 1e-14 || c2 > 1e-14 || c3 > 2e-6 || c4 > 2e-6) bad = 1;
    /* cached and table twiddles give the same bits */
    if (c1 != e1 || c2 != e2) { printf("cached twiddles differ from table twiddles at N=%d\n", N); bad = 1; }
    return bad;
}

/* usage: fft_emul [max_n] (default 4096: every size) */
int main(int argc, char **argv) {
    const int maxn = argc > 1 ? atoi(argv[1]) : 4096;
    int bad = 0;
    if (maxn >= 512) bad |= run_t<512>();
    if (maxn >= 1024) bad |= run_t<1024>();
    if (maxn >= 2048) bad |= run_t<2048>();
    if (maxn >= 4096) bad |= run_t<4096>();
    bad |= run_n<16>(); bad |= run_n<32>(); bad |= run_n<64>(); bad |= run_n<128>();
    bad |= run_n<192>(); bad |= run_n<256>();
    if (maxn >= 512) bad |= run_n<512>();
    if (maxn >= 1024) bad |= run_n<1024>();
    if (maxn >= 2048) bad |= run_n<2048>();
    if (maxn >= 4096) bad |= run_n<4096>();
    printf(bad ? "FAIL\n" : "OK\n");
    return bad;
}
