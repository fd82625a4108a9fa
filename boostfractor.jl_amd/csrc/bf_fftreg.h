// bf_fftreg.h -- register-resident Stockham FFT for power-of-two lines.
//
// A line of N = 2^k points is transformed by T = N/E threads, each holding E
// points in registers in "natural" layout: register m of thread t holds point
// x = t + T*m.  The transform is a sequence of Stockham autosort passes
// (radix E, ..., radix E, then one smaller radix if E does not divide the
// remaining length); the first pass reads the natural layout straight from
// registers and the last pass produces the natural layout again, so only the
// exchanges between passes touch LDS (one exchange for N <= E^2).  Global
// loads and stores of a line in natural layout are coalesced (consecutive
// threads own consecutive points), and pointwise work between an inverse and a
// forward transform stays in registers.
//
// Pass p (radix R, stride Ns = product of earlier radices) computes, for every
// virtual thread j in [0, N/R):  k = j mod Ns,
//     a[r] = x[j + r N/R] * w_N^(r k N/(Ns R)),  a = DFT_R(a),
//     y[(j - k) R + k + r Ns] = a[r].
// Small DFTs use the split radix-4 decomposition with exact constant twiddles.
#pragma once
#include "bf_device.h"

namespace bf {

// exp(-2 pi i m/16), m = 0..15
__device__ constexpr double W16R[16] = {1.0,
                                        0.92387953251128673848,
                                        0.70710678118654752440,
                                        0.38268343236508977173,
                                        0.0,
                                        -0.38268343236508977173,
                                        -0.70710678118654752440,
                                        -0.92387953251128673848,
                                        -1.0,
                                        -0.92387953251128673848,
                                        -0.70710678118654752440,
                                        -0.38268343236508977173,
                                        0.0,
                                        0.38268343236508977173,
                                        0.70710678118654752440,
                                        0.92387953251128673848};
__device__ constexpr double W16I[16] = {0.0,
                                        -0.38268343236508977173,
                                        -0.70710678118654752440,
                                        -0.92387953251128673848,
                                        -1.0,
                                        -0.92387953251128673848,
                                        -0.70710678118654752440,
                                        -0.38268343236508977173,
                                        0.0,
                                        0.38268343236508977173,
                                        0.70710678118654752440,
                                        0.92387953251128673848,
                                        1.0,
                                        0.92387953251128673848,
                                        0.70710678118654752440,
                                        0.38268343236508977173};

// multiply by w_R^m (R | 16), exact for the trivial angles
template <typename T, int R, bool INV>
__device__ __forceinline__ cplx<T> twc(cplx<T> a, int m) {
    const int q = (m * (16 / R)) & 15;
    if (q == 0) return a;
    T wr = (T)W16R[q], wi = (T)W16I[q];
    if (INV) wi = -wi;
    if (q == 4) return INV ? cmk<T>(-a.im, a.re) : cmk<T>(a.im, -a.re);
    if (q == 8) return cmk<T>(-a.re, -a.im);
    if (q == 12) return INV ? cmk<T>(a.im, -a.re) : cmk<T>(-a.im, a.re);
    return cmk<T>(a.re * wr - a.im * wi, a.re * wi + a.im * wr);
}

template <typename T, bool INV>
__device__ __forceinline__ void dft2(cplx<T>& a0, cplx<T>& a1) {
    const cplx<T> t = a0;
    a0 = cadd(t, a1);
    a1 = csub(t, a1);
}

// in-register DFT of R points (R = 2, 4, 8, 16), natural order in and out
template <typename T, int R, bool INV>
__device__ __forceinline__ void dft_reg(cplx<T>* a) {
    if constexpr (R == 2) {
        dft2<T, INV>(a[0], a[1]);
    } else if constexpr (R == 4) {
        radix4<T, INV>(a[0], a[1], a[2], a[3]);
    } else {
        // R = 4 Q: n = Q n1 + n2, k = k1 + 4 k2
        constexpr int Q = R / 4;
        cplx<T> b[R];
#pragma unroll
        for (int n2 = 0; n2 < Q; n2++) {
            cplx<T> c0 = a[n2], c1 = a[Q + n2], c2 = a[2 * Q + n2], c3 = a[3 * Q + n2];
            radix4<T, INV>(c0, c1, c2, c3);
            b[n2 * 4 + 0] = c0;
            b[n2 * 4 + 1] = twc<T, R, INV>(c1, n2 * 1);
            b[n2 * 4 + 2] = twc<T, R, INV>(c2, n2 * 2);
            b[n2 * 4 + 3] = twc<T, R, INV>(c3, n2 * 3);
        }
#pragma unroll
        for (int k1 = 0; k1 < 4; k1++) {
            cplx<T> d[Q];
#pragma unroll
            for (int n2 = 0; n2 < Q; n2++) d[n2] = b[n2 * 4 + k1];
            if constexpr (Q == 2)
                dft2<T, INV>(d[0], d[1]);
            else
                radix4<T, INV>(d[0], d[1], d[2], d[3]);
#pragma unroll
            for (int k2 = 0; k2 < Q; k2++) a[k1 + 4 * k2] = d[k2];
        }
    }
}

// padded LDS index: one pad slot per E points keeps the pass-1 stores conflict free
template <int E>
__device__ __forceinline__ int lpad(int i) {
    return i + i / E;
}
template <int N, int E>
struct RegFFT {
    static constexpr int T = N / E;                 // threads per line
    static constexpr int LDS = N + N / E;           // padded line length (elements)
};

// One Stockham pass over LDS data (not the first pass): radix R at stride NS.
// Reads x from `s` (padded), writes y either to `s` (after a barrier) or, for
// the last pass, into registers in natural layout.
template <typename T, typename L, int N, int E, int R, int NS, bool INV, bool LAST>
__device__ __forceinline__ void reg_pass(cplx<T> (&v)[E], cplx<L>* s, int t, const cplx<T>* __restrict__ tw) {
    constexpr int TT = N / E;
    constexpr int V = E / R;  // virtual threads per thread
    cplx<T> a[V][R];
#pragma unroll
    for (int vv = 0; vv < V; vv++) {
        const int j = t + vv * TT;
        const int k = j & (NS - 1);
#pragma unroll
        for (int r = 0; r < R; r++) a[vv][r] = ccast<T>(s[lpad<E>(j + r * (N / R))]);
        if constexpr (sizeof(T) == 8) {
            // FP64: w^r by recurrence from one table entry (keeps register pressure low;
            // the recurrence adds at most ~R ulp to a twiddle)
            const cplx<T> w1 = tw_get<T, INV>(tw, (k * (N / (NS * R))) & (N - 1));
            cplx<T> w = w1;
#pragma unroll
            for (int r = 1; r < R; r++) {
                a[vv][r] = cmul(a[vv][r], w);
                if (r + 1 < R) w = cmul(w, w1);
            }
        } else {
#pragma unroll
            for (int r = 1; r < R; r++) {
                const int m = (r * k * (N / (NS * R))) & (N - 1);
                a[vv][r] = cmul(a[vv][r], tw_get<T, INV>(tw, m));
            }
        }
        dft_reg<T, R, INV>(a[vv]);
    }
    if constexpr (LAST) {
        // j < N/R = NS here: output point j + r NS = t + T (vv + r V)
#pragma unroll
        for (int vv = 0; vv < V; vv++)
#pragma unroll
            for (int r = 0; r < R; r++) v[vv + r * V] = a[vv][r];
    } else {
        __syncthreads();
#pragma unroll
        for (int vv = 0; vv < V; vv++) {
            const int j = t + vv * TT;
            const int k = j & (NS - 1);
#pragma unroll
            for (int r = 0; r < R; r++) s[lpad<E>((j - k) * R + k + r * NS)] = ccast<L>(a[vv][r]);
        }
        __syncthreads();
    }
}

// Full transform of one line held in registers (natural layout in and out).
// `s` is this line's LDS buffer of RegFFT<N,E>::LDS elements of type L: the
// exchanges between passes are held in L (the field storage precision) while
// all arithmetic runs in T.  Every thread of the block must call it (barriers).
template <typename T, typename L, int N, int E, bool INV>
__device__ __forceinline__ void fft_reg(cplx<T> (&v)[E], cplx<L>* s, int t, const cplx<T>* __restrict__ tw) {
    constexpr int TT = N / E;
    if constexpr (N == E) {
        dft_reg<T, E, INV>(v);
    } else {
        // pass 1: radix E at stride 1 straight from registers (j = t, k = 0)
        dft_reg<T, E, INV>(v);
        __syncthreads();  // the buffer may still be read by a previous transform
#pragma unroll
        for (int r = 0; r < E; r++) s[lpad<E>(t * E + r)] = ccast<L>(v[r]);
        __syncthreads();
        // remaining passes
        constexpr int N2 = N / E;  // length still to combine
        if constexpr (N2 <= E) {
            reg_pass<T, L, N, E, N2, E, INV, true>(v, s, t, tw);
        } else {
            constexpr int N3 = N2 / E;
            if constexpr (N3 <= E && N3 >= 2) {
                reg_pass<T, L, N, E, E, E, INV, false>(v, s, t, tw);
                reg_pass<T, L, N, E, N3, E * E, INV, true>(v, s, t, tw);
            } else if constexpr (N3 == 1) {
                reg_pass<T, L, N, E, E, E, INV, true>(v, s, t, tw);
            } else {
                static_assert(N3 <= E, "line too long for the register FFT");
            }
        }
        (void)TT;
    }
}

}  // namespace bf
