// Host harness for the band-pruned register FFT of
// boostfractor.jl_amd/csrc/bf_fftreg.h.  The test compiles a copy of that
// header with its device qualifiers removed (FFTREG_HOST names the copy), runs
// every thread of a line as a pthread (a pthread barrier stands in for
// __syncthreads) and checks
//   * the masked transforms (fft_reg_m) against the unmasked ones (fft_reg),
//     bit for bit, on the elements the output mask keeps, with inputs zero
//     outside the input mask;
//   * the whole-field completion (fft_reg_m with the band output mask, then
//     fft_reg_last_m with the complement) against fft_reg on every element;
//   * fft_reg against a direct DFT.
// Prints one line per case and "total mismatches <n> max_dft_err <e>".
#include <pthread.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <random>
#include <type_traits>
#include <vector>

template <typename T>
struct alignas(2 * sizeof(T)) cplx {
    T re, im;
};
template <typename T>
inline cplx<T> cmk(T a, T b) {
    cplx<T> z;
    z.re = a;
    z.im = b;
    return z;
}
template <typename T>
inline cplx<T> cadd(cplx<T> a, cplx<T> b) { return cmk<T>(a.re + b.re, a.im + b.im); }
template <typename T>
inline cplx<T> csub(cplx<T> a, cplx<T> b) { return cmk<T>(a.re - b.re, a.im - b.im); }
template <typename T>
inline cplx<T> cmul(cplx<T> a, cplx<T> b) { return cmk<T>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
template <typename To, typename Ti>
inline cplx<To> ccast(cplx<Ti> a) { return cmk<To>((To)a.re, (To)a.im); }
template <typename T, bool INV>
inline cplx<T> tw_get(const cplx<T>* tw, int m) {
    cplx<T> w = tw[m];
    if (INV) w.im = -w.im;
    return w;
}
// radix-4 butterfly with the operation order of the device header's radix4
template <typename T, bool INV>
inline void radix4(cplx<T>& a0, cplx<T>& a1, cplx<T>& a2, cplx<T>& a3) {
    cplx<T> t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = csub(a1, a3);
    if (!INV)
        t3 = cmk<T>(t3.im, -t3.re);
    else
        t3 = cmk<T>(-t3.im, t3.re);
    a0 = cadd(t0, t2);
    a2 = csub(t0, t2);
    a1 = cadd(t1, t3);
    a3 = csub(t1, t3);
}
static pthread_barrier_t bar;
#define __syncthreads() pthread_barrier_wait(&bar)
#include FFTREG_HOST
using namespace bf;

enum Mode { FULL, MASKED, COMPLETED };

template <int N, bool INV, int IM, int OM, Mode MODE>
struct Line {
    static constexpr int E = 16, TT = N / E, LD = N + N / E;
    std::vector<cplx<double>> in, out, tw, lds;
    static void* run(void* p) {
        auto* a = (std::pair<Line*, int>*)p;
        Line* j = a->first;
        const int t = a->second;
        cplx<double> v[E];
        for (int m = 0; m < E; m++) v[m] = j->in[t + TT * m];
        if (MODE == FULL) {
            fft_reg<double, double, N, E, INV>(v, j->lds.data(), t, j->tw.data());
        } else {
            fft_reg_m<double, double, N, E, INV, IM, OM>(v, j->lds.data(), t, j->tw.data());
            if (MODE == COMPLETED) fft_reg_last_m<double, double, N, E, INV, 0xFFFF & ~OM>(v, j->lds.data(), t, j->tw.data());
        }
        const int keep = MODE == MASKED ? OM : 0xFFFF;
        for (int m = 0; m < E; m++)
            if ((keep >> m) & 1) j->out[t + TT * m] = v[m];
        return nullptr;
    }
    void go() {
        lds.assign(LD, cmk(0.0, 0.0));
        out.assign(N, cmk<double>(NAN, NAN));
        pthread_barrier_init(&bar, nullptr, TT);
        std::vector<pthread_t> th(TT);
        std::vector<std::pair<Line*, int>> args(TT);
        for (int t = 0; t < TT; t++) {
            args[t] = {this, t};
            pthread_create(&th[t], nullptr, run, &args[t]);
        }
        for (int t = 0; t < TT; t++) pthread_join(th[t], nullptr);
        pthread_barrier_destroy(&bar);
    }
};

static double max_dft_err = 0.0;

template <int N, bool INV, int IM, int OM>
int check(unsigned seed) {
    constexpr int TT = N / 16;
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> nd;
    std::vector<cplx<double>> tw(N), in(N);
    for (int m = 0; m < N; m++) tw[m] = cmk(std::cos(-2 * M_PI * m / N), std::sin(-2 * M_PI * m / N));
    for (int x = 0; x < N; x++) in[x] = ((IM >> (x / TT)) & 1) ? cmk(nd(rng), nd(rng)) : cmk(0.0, 0.0);
    Line<N, INV, 0xFFFF, 0xFFFF, FULL> full;
    Line<N, INV, IM, OM, MASKED> masked;
    Line<N, INV, IM, OM, COMPLETED> done;
    full.in = masked.in = done.in = in;
    full.tw = masked.tw = done.tw = tw;
    full.go();
    masked.go();
    done.go();
    int bad = 0;
    double dft = 0.0;
    for (int x = 0; x < N; x++) {
        std::complex<double> s = 0;
        for (int y = 0; y < N; y++)
            s += std::complex<double>(in[y].re, in[y].im) *
                 std::polar(1.0, (INV ? 2 : -2) * M_PI * (double)((long)x * y % N) / N);
        dft = std::max(dft, std::abs(s - std::complex<double>(full.out[x].re, full.out[x].im)) / std::sqrt((double)N));
        const bool kept = (OM >> (x / TT)) & 1;
        if (kept && (masked.out[x].re != full.out[x].re || masked.out[x].im != full.out[x].im)) bad++;
        if (done.out[x].re != full.out[x].re || done.out[x].im != full.out[x].im) bad++;
    }
    max_dft_err = std::max(max_dft_err, dft);
    std::printf("N=%d INV=%d IM=%04x OM=%04x mismatches %d dft_err %.2e\n", N, (int)INV, IM, OM, bad, dft);
    return bad;
}

int main() {
    int bad = 0;
    bad += check<128, false, 0x07E0, 0xFFFF>(1);
    bad += check<128, true, 0xFFFF, 0x07E0>(2);
    bad += check<256, false, 0x07E0, 0xFFFF>(3);
    bad += check<256, true, 0xFFFF, 0x07E0>(4);
    bad += check<256, true, 0x07E0, 0x07E0>(5);
    bad += check<512, false, 0x07E0, 0xFFFF>(6);
    bad += check<512, true, 0xFFFF, 0x07E0>(7);
    bad += check<1024, false, 0x1FF8, 0xFFFF>(8);
    bad += check<1024, true, 0xFFFF, 0x1FF8>(9);
    bad += check<128, false, 0x1FF8, 0xFFFF>(10);
    bad += check<128, true, 0xFFFF, 0x1FF8>(11);
    bad += check<256, true, 0xFFFF, 0x1FF8>(12);
    std::printf("total mismatches %d max_dft_err %.3e\n", bad, max_dft_err);
    return bad != 0;
}
