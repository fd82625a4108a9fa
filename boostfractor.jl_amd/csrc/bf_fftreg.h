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
#include <type_traits>

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


// ---------------------------------------------------------------------------
// Band-pruned transforms.  Bit m of IM set: input point m of the thread may be
// nonzero (the others are exact zeros and are not read); bit m of OM set:
// output point m is needed (the others are left unset).  Used where a line is
// known to be zero outside a band of rows (or columns) and only that band of
// the result is kept: the first pass skips the butterflies of zero inputs and
// the last pass skips the outputs nobody reads.  In IEEE arithmetic every value
// that is computed equals the unpruned transform's (x + 0 = x and 0 - x = -x
// exactly; only the sign of an exact zero can differ); with FMA contraction the
// compiler may fuse a product into a sum where the pruned code uses it once and
// the unpruned code twice, which moves results by an ulp.
// ---------------------------------------------------------------------------
template <int I, int N, typename F>
__device__ __forceinline__ void sfor(F&& f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        sfor<I + 1, N>(f);
    }
}

template <typename T, bool ZA, bool ZB>
__device__ __forceinline__ cplx<T> zadd(const cplx<T>& a, const cplx<T>& b) {
    if constexpr (ZA && ZB) return cmk<T>(0, 0);
    else if constexpr (ZA) return b;
    else if constexpr (ZB) return a;
    else return cadd(a, b);
}
template <typename T, bool ZA, bool ZB>
__device__ __forceinline__ cplx<T> zsub(const cplx<T>& a, const cplx<T>& b) {
    if constexpr (ZA && ZB) return cmk<T>(0, 0);
    else if constexpr (ZB) return a;
    else if constexpr (ZA) return cmk<T>(-b.re, -b.im);
    else return csub(a, b);
}

template <typename T, int IM, int OM>
__device__ __forceinline__ void dft2m(cplx<T>& a0, cplx<T>& a1) {
    constexpr bool z0 = !(IM & 1), z1 = !(IM & 2);
    const cplx<T> t = a0;
    if constexpr ((OM & 1) != 0) a0 = zadd<T, z0, z1>(t, a1);
    if constexpr ((OM & 2) != 0) a1 = zsub<T, z0, z1>(t, a1);
}

// radix4 with 4-bit input/output masks (same operation order as radix4)
template <typename T, bool INV, int IM, int OM>
__device__ __forceinline__ void radix4m(cplx<T>& a0, cplx<T>& a1, cplx<T>& a2, cplx<T>& a3) {
    constexpr bool z0 = !(IM & 1), z1 = !(IM & 2), z2 = !(IM & 4), z3 = !(IM & 8);
    constexpr bool zt0 = z0 && z2, zt2 = z1 && z3;  // t1 is zero with t0, t3 with t2
    const cplx<T> t0 = zadd<T, z0, z2>(a0, a2), t1 = zsub<T, z0, z2>(a0, a2);
    const cplx<T> t2 = zadd<T, z1, z3>(a1, a3);
    cplx<T> t3 = zsub<T, z1, z3>(a1, a3);
    if (!INV)
        t3 = cmk<T>(t3.im, -t3.re);
    else
        t3 = cmk<T>(-t3.im, t3.re);
    if constexpr ((OM & 1) != 0) a0 = zadd<T, zt0, zt2>(t0, t2);
    if constexpr ((OM & 4) != 0) a2 = zsub<T, zt0, zt2>(t0, t2);
    if constexpr ((OM & 2) != 0) a1 = zadd<T, zt0, zt2>(t1, t3);
    if constexpr ((OM & 8) != 0) a3 = zsub<T, zt0, zt2>(t1, t3);
}

// masks of the split radix-4 decomposition R = 4 Q (n = Q n1 + n2, k = k1 + 4 k2)
__host__ __device__ constexpr int group_in(int IM, int Q, int n2) {  // inputs a[j Q + n2], j = 0..3
    return ((IM >> n2) & 1) | (((IM >> (Q + n2)) & 1) << 1) | (((IM >> (2 * Q + n2)) & 1) << 2) |
           (((IM >> (3 * Q + n2)) & 1) << 3);
}
__host__ __device__ constexpr int stage2_in(int IM, int Q) {  // bit n2: group n2 has an input
    int m = 0;
    for (int n2 = 0; n2 < Q; n2++)
        if (group_in(IM, Q, n2) != 0) m |= 1 << n2;
    return m;
}
__host__ __device__ constexpr int stage2_out(int OM, int Q, int k1) {  // bit k2: output k1 + 4 k2 needed
    int m = 0;
    for (int k2 = 0; k2 < Q; k2++)
        if ((OM >> (k1 + 4 * k2)) & 1) m |= 1 << k2;
    return m;
}
__host__ __device__ constexpr int stage1_out(int OM, int Q) {  // bit k1: some output k1 + 4 k2 needed
    int m = 0;
    for (int k1 = 0; k1 < 4; k1++)
        if (stage2_out(OM, Q, k1) != 0) m |= 1 << k1;
    return m;
}

// dft_reg with masks (same operation order as dft_reg)
template <typename T, int R, bool INV, int IM, int OM>
__device__ __forceinline__ void dft_reg_m(cplx<T>* a) {
    if constexpr (R == 2) {
        dft2m<T, IM & 3, OM & 3>(a[0], a[1]);
    } else if constexpr (R == 4) {
        radix4m<T, INV, IM & 15, OM & 15>(a[0], a[1], a[2], a[3]);
    } else {
        constexpr int Q = R / 4;
        constexpr int OM1 = stage1_out(OM, Q);
        constexpr int IM2 = stage2_in(IM, Q);
        cplx<T> b[R];
        sfor<0, Q>([&](auto n2c) {
            constexpr int n2 = decltype(n2c)::value;
            constexpr int gi = group_in(IM, Q, n2);
            if constexpr (gi != 0) {
                const cplx<T> zero = cmk<T>(0, 0);
                cplx<T> c0 = (gi & 1) ? a[n2] : zero, c1 = (gi & 2) ? a[Q + n2] : zero;
                cplx<T> c2 = (gi & 4) ? a[2 * Q + n2] : zero, c3 = (gi & 8) ? a[3 * Q + n2] : zero;
                radix4m<T, INV, gi, OM1>(c0, c1, c2, c3);
                if constexpr ((OM1 & 1) != 0) b[n2 * 4 + 0] = c0;
                if constexpr ((OM1 & 2) != 0) b[n2 * 4 + 1] = twc<T, R, INV>(c1, n2 * 1);
                if constexpr ((OM1 & 4) != 0) b[n2 * 4 + 2] = twc<T, R, INV>(c2, n2 * 2);
                if constexpr ((OM1 & 8) != 0) b[n2 * 4 + 3] = twc<T, R, INV>(c3, n2 * 3);
            }
        });
        sfor<0, 4>([&](auto k1c) {
            constexpr int k1 = decltype(k1c)::value;
            constexpr int om2 = stage2_out(OM, Q, k1);
            if constexpr (om2 != 0) {
                cplx<T> d[Q];
                sfor<0, Q>([&](auto n2c) {
                    constexpr int n2 = decltype(n2c)::value;
                    if constexpr (((IM2 >> n2) & 1) != 0)
                        d[n2] = b[n2 * 4 + k1];
                    else
                        d[n2] = cmk<T>(0, 0);
                });
                if constexpr (Q == 2)
                    dft2m<T, IM2, om2>(d[0], d[1]);
                else
                    radix4m<T, INV, IM2, om2>(d[0], d[1], d[2], d[3]);
                sfor<0, Q>([&](auto k2c) {
                    constexpr int k2 = decltype(k2c)::value;
                    if constexpr (((om2 >> k2) & 1) != 0) a[k1 + 4 * k2] = d[k2];
                });
            }
        });
    }
}

// bit r: output point vv + r V of the last pass is needed
__host__ __device__ constexpr int last_out(int OM, int vv, int V, int R) {
    int m = 0;
    for (int r = 0; r < R; r++)
        if ((OM >> (vv + r * V)) & 1) m |= 1 << r;
    return m;
}

// the last Stockham pass of reg_pass with an output mask: virtual threads whose
// outputs are all unused are skipped whole
template <typename T, typename L, int N, int E, int R, int NS, bool INV, int OM>
__device__ __forceinline__ void reg_pass_last_m(cplx<T> (&v)[E], cplx<L>* s, int t, const cplx<T>* __restrict__ tw) {
    constexpr int TT = N / E;
    constexpr int V = E / R;
    sfor<0, V>([&](auto vvc) {
        constexpr int vv = decltype(vvc)::value;
        constexpr int om = last_out(OM, vv, V, R);
        if constexpr (om != 0) {
            cplx<T> a[R];
            const int j = t + vv * TT;
            const int k = j & (NS - 1);
#pragma unroll
            for (int r = 0; r < R; r++) a[r] = ccast<T>(s[lpad<E>(j + r * (N / R))]);
            if constexpr (sizeof(T) == 8) {
                const cplx<T> w1 = tw_get<T, INV>(tw, (k * (N / (NS * R))) & (N - 1));
                cplx<T> w = w1;
#pragma unroll
                for (int r = 1; r < R; r++) {
                    a[r] = cmul(a[r], w);
                    if (r + 1 < R) w = cmul(w, w1);
                }
            } else {
#pragma unroll
                for (int r = 1; r < R; r++) {
                    const int m = (r * k * (N / (NS * R))) & (N - 1);
                    a[r] = cmul(a[r], tw_get<T, INV>(tw, m));
                }
            }
            dft_reg_m<T, R, INV, (1 << R) - 1, om>(a);
#pragma unroll
            for (int r = 0; r < R; r++)
                if ((om >> r) & 1) v[vv + r * V] = a[r];
        }
    });
}

// fft_reg with an input mask on the first pass and an output mask on the last
template <typename T, typename L, int N, int E, bool INV, int IM, int OM>
__device__ __forceinline__ void fft_reg_m(cplx<T> (&v)[E], cplx<L>* s, int t, const cplx<T>* __restrict__ tw) {
    static_assert(E == 16, "masks are 16-bit");
    if constexpr (N == E) {
        dft_reg_m<T, E, INV, IM, OM>(v);
    } else {
        dft_reg_m<T, E, INV, IM, 0xFFFF>(v);
        __syncthreads();  // the buffer may still be read by a previous transform
#pragma unroll
        for (int r = 0; r < E; r++) s[lpad<E>(t * E + r)] = ccast<L>(v[r]);
        __syncthreads();
        constexpr int N2 = N / E;
        if constexpr (N2 <= E) {
            reg_pass_last_m<T, L, N, E, N2, E, INV, OM>(v, s, t, tw);
        } else {
            constexpr int N3 = N2 / E;
            static_assert(N3 <= E, "line too long for the register FFT");
            if constexpr (N3 == 1) {
                reg_pass_last_m<T, L, N, E, E, E, INV, OM>(v, s, t, tw);
            } else {
                reg_pass<T, L, N, E, E, E, INV, false>(v, s, t, tw);
                reg_pass_last_m<T, L, N, E, N3, E * E, INV, OM>(v, s, t, tw);
            }
        }
    }
}

// The last pass of fft_reg_m again, with another output mask, for a transform
// whose last-pass input is still in `s` (nothing has written `s` since).
template <typename T, typename L, int N, int E, bool INV, int OM>
__device__ __forceinline__ void fft_reg_last_m(cplx<T> (&v)[E], cplx<L>* s, int t, const cplx<T>* __restrict__ tw) {
    static_assert(N > E, "the single-pass transform keeps no input in LDS");
    constexpr int N2 = N / E;
    if constexpr (N2 <= E)
        reg_pass_last_m<T, L, N, E, N2, E, INV, OM>(v, s, t, tw);
    else
        reg_pass_last_m<T, L, N, E, N2 / E, E * E, INV, OM>(v, s, t, tw);
}

// elements m in [BM, 15 - BM] of a 16-element thread
__host__ __device__ constexpr int band_mask16(int BM) { return BM <= 0 ? 0xFFFF : ((1 << (16 - 2 * BM)) - 1) << BM; }

}  // namespace bf
