// bf_transformer.hip -- mode matching (src/boost_fractor_transformer.jl) on gfx950.
//
// Every propagation matrix of the reference,
//     P[m,m'] = sum_c sum_x conj(Phi_m,c(x)) * prop(Phi_m',c)(x)          (:195-210)
// is evaluated as a weighted Gram product P = A^H diag(w) B on the FP64 matrix
// cores (v_mfma_f64_16x16x4f64):
//   * air region, flat and untilted (propagator / propagatorNoTilts): by
//     Parseval, A = FFT(Phi), B = FFT(mask*Phi), w = exp(-i kz dz)/K -- the
//     frequency-independent transforms are computed once per call;
//   * disk region: A = B = Phi, w = tilt*surface phase (or no weight at all
//     for flat untilted disks: the Gram matrix is shared by every frequency);
//     then the left multiplication by diag(exp(-i kz_m dz)) (:212-221);
//   * anything else (tilted/rough air, momentum-space or 1D propagators): the
//     patterns are propagated explicitly on the device and A = Phi, B = prop(Phi).
// The reduction over the grid is split across blocks into partial matrices that
// a second kernel sums in a fixed order (deterministic results).
//
// The 2n x 2n transfer-matrix chain (:326-341) and the final solves (:346,:356)
// stay in FP64: batched LU with partial pivoting (pivot = max |re|+|im|, as
// LAPACK's izamax) for inv(P) and for T22 x = b, and one block per frequency
// walking the regions, using the block structure G*[P 0; 0 inv(P)] =
// [a P, b inv(P); b P, a inv(P)].
#include "bf_modematch.h"

#include <chrono>
#include <map>
#include <memory>
#include <mutex>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace bf {

// ---------------------------------------------------------------------------
// Bessel modes (src/boost_fractor_transformer.jl:88-107)
// ---------------------------------------------------------------------------
// J_n(x) for integer n by Miller's backward recurrence (J.C.P. Miller, 1952),
// normalised with J_0 + 2 sum_k J_2k = 1; J_-n = (-1)^n J_n.  Starting order
// max(n, x) + 30 + sqrt(60 max(n, x)) keeps the absolute error ~1e-16 for the
// orders and arguments of mode patterns.
__host__ __device__ double besselj_dev(int n, double x) {
    int sgn = 1;
    if (n < 0) {
        n = -n;
        if (n & 1) sgn = -sgn;
    }
    if (x < 0) {
        x = -x;
        if (n & 1) sgn = -sgn;
    }
    if (x == 0.0) return n == 0 ? 1.0 : 0.0;
    const double big = n > x ? (double)n : x;
    const int m = 2 * (int)((big + 30.0 + sqrt(60.0 * big)) / 2.0) + 2;
    double jp = 0.0, j = 1e-30, sum = 0.0, ans = 0.0;
    for (int k = m; k > 0; k--) {
        const double jm = (2.0 * k / x) * j - jp;  // J_{k-1}
        jp = j;
        j = jm;
        if (fabs(j) > 1e250) {
            j *= 1e-250;
            jp *= 1e-250;
            ans *= 1e-250;
            sum *= 1e-250;
        }
        if (k - 1 == n) ans = j;
        if (k - 1 >= 2 && ((k - 1) & 1) == 0) sum += j;
    }
    sum = 2.0 * sum + j;
    return sgn * ans / sum;
}

// m-th positive zero of J_nu: sign scan (zero spacing of J_nu exceeds pi for
// nu >= 1/2, and j_{0,1} > 2) then bisection to full precision.
__device__ double besselj_zero_dev(int nu, int m) {
    double x = (nu == 0) ? 0.5 : (double)nu;
    const double h = 0.25;
    double fa = besselj_dev(nu, x);
    int found = 0;
    for (int it = 0; it < 100000; it++) {
        const double xb = x + h, fb = besselj_dev(nu, xb);
        if ((fa < 0.0 && fb >= 0.0) || (fa > 0.0 && fb <= 0.0)) {
            if (++found == m) {
                double lo = x, hi = xb, flo = fa;
                for (int k = 0; k < 200; k++) {
                    const double mid = 0.5 * (lo + hi);
                    if (mid <= lo || mid >= hi) break;
                    const double fm = besselj_dev(nu, mid);
                    if (fm == 0.0) { lo = hi = mid; break; }
                    if ((flo < 0.0) == (fm < 0.0)) { lo = mid; flo = fm; } else { hi = mid; }
                }
                return fabs(besselj_dev(nu, lo)) <= fabs(besselj_dev(nu, hi)) ? lo : hi;
            }
        }
        x = xb;
        fa = fb;
    }
    return NAN;
}

__global__ void k_mode_kt(zd* kt, int M, int L, double diskR) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = M * (2 * L + 1);
    if (idx >= n) return;
    const int m = idx / (2 * L + 1) + 1, l = idx % (2 * L + 1) - L;
    kt[idx] = cmk<double>(besselj_zero_dev(l < 0 ? -l : l, m) / diskR, 0.0);
}

// pattern = besselj(l, kr*RR) * exp(-1im*l*Phi), zero where RR > diskR
__global__ void k_mode_patterns(zd* pat, const zd* __restrict__ kt, int L, int nx, int ny,
                                const double* __restrict__ X, const double* __restrict__ Y, double diskR) {
    const int idx = blockIdx.y;
    const int l = idx % (2 * L + 1) - L;
    const double kr = kt[idx].re;
    const size_t K = (size_t)nx * ny;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < K; i += (size_t)gridDim.x * blockDim.x) {
        const double x = X[i % nx], y = Y[i / nx];
        const double RR = sqrt(x * x + y * y);
        zd v = cmk<double>(0.0, 0.0);
        if (!(RR > diskR)) {
            const double phi = atan2(y, x);
            const double j = besselj_dev(l, kr * RR);
            const zd e = cexp_j(cmk<double>(phi * (0.0 * l), phi * (-1.0 * l)));
            v = cmk<double>(j * e.re, j * e.im);
        }
        pat[(size_t)idx * K + i] = v;
    }
}

// out[b] = sum |v|^2 over n elements of row b (fixed-order two-level reduction)
__global__ void k_row_norm2(const zd* __restrict__ v, size_t n, double* out) {
    __shared__ double red[16];
    const zd* r = v + (size_t)blockIdx.x * n;
    double s = 0.0;
    for (size_t i = threadIdx.x; i < n; i += blockDim.x) s += r[i].re * r[i].re + r[i].im * r[i].im;
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += red[w];
        out[blockIdx.x] = t;
    }
}

__global__ void k_row_divide_sqrt(zd* v, size_t n, const double* __restrict__ norm2) {
    const double d = sqrt(norm2[blockIdx.y]);
    zd* r = v + (size_t)blockIdx.y * n;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        r[i] = cmk<double>(r[i].re / d, r[i].im / d);
}

// axion_induced_modes source: B (or ones) * velocity phase * [sqrt(x^2+y^2) <= diskR]
__global__ void k_axion_source(zd* b, const zd* __restrict__ B, int ncomp, int nx, int ny,
                               const double* __restrict__ X, const double* __restrict__ Y, double diskR, double vel,
                               double Ma) {
    const size_t K = (size_t)nx * ny;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < K * ncomp; i += (size_t)gridDim.x * blockDim.x) {
        const size_t p = i % K;
        const double x = X[p % nx], y = Y[p / nx];
        zd v = B ? B[i] : cmk<double>(1.0, 0.0);
        if (vel != 0.0) {
            // exp(-1im*Ma*(-v)*x)
            const zd a = cmk<double>(Ma * 0.0, Ma * -1.0);
            const zd e = cexp_j(rmul(x, rmul(-vel, a)));
            v = cmul(v, e);
        }
        if (!(sqrt(x * x + y * y) <= diskR)) v = cmk<double>(0.0 * v.re, 0.0 * v.im);
        b[i] = v;
    }
}

// out[m] = sum_k conj(pat[m][k]) * b[k]   (one block per mode)
__global__ void k_project(const zd* __restrict__ pat, const zd* __restrict__ b, size_t n, zd* out) {
    __shared__ double red[32];
    const zd* r = pat + (size_t)blockIdx.x * n;
    double sr = 0.0, si = 0.0;
    for (size_t i = threadIdx.x; i < n; i += blockDim.x) {
        const zd v = cmulc(b[i], r[i]);
        sr += v.re;
        si += v.im;
    }
    for (int o = 32; o > 0; o >>= 1) {
        sr += __shfl_xor(sr, o);
        si += __shfl_xor(si, o);
    }
    if ((threadIdx.x & 63) == 0) {
        red[2 * (threadIdx.x >> 6)] = sr;
        red[2 * (threadIdx.x >> 6) + 1] = si;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tr = 0.0, ti = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
            tr += red[2 * w];
            ti += red[2 * w + 1];
        }
        out[blockIdx.x] = cmk<double>(tr, ti);
    }
}

// modes2field: out[k] = sum_m c[m] * pat[m][k]   (in mode order)
__global__ void k_modes2field(const zd* __restrict__ pat, const zd* __restrict__ c, int n, size_t len, zd* out) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (size_t)gridDim.x * blockDim.x) {
        zd s = cmk<double>(0.0, 0.0);
        for (int m = 0; m < n; m++) s = cadd(s, cmul(c[m], pat[(size_t)m * len + i]));
        out[i] = s;
    }
}

// ---------------------------------------------------------------------------
// weighted Gram products on FP64 MFMA: P_p = A_p^H diag(w_p) B_p
// ---------------------------------------------------------------------------
struct GemmProb {
    const zd* A;  // [n][Ktot]
    const zd* B;  // [n][Ktot]
    const zd* w;  // [Kw] weights applied at index k % Kw, or null
    int out;      // output slot
    int pad;
};

typedef double __attribute__((ext_vector_type(4))) d4;

constexpr int KC = 32;          // k values staged per chunk
constexpr int LDA = KC + 1;     // padded LDS row (complex elements)
constexpr int TILES_PER_WAVE = 2;  // 2 tiles x 4 accumulators x 8 registers: no spills at 1024 threads

// grid: (splits, problems, tile batches); block: 64 * waves; batch z covers the
// output tiles [z * waves * TILES_PER_WAVE, ...).  Output partial
// [prob][split][n][n] col-major.
__global__ void __launch_bounds__(1024) k_wgram(const GemmProb* __restrict__ probs, int n, size_t Ktot, size_t Kw,
                                                 size_t kper, zd* partial) {
    extern __shared__ unsigned char smem_raw[];
    const GemmProb pr = probs[blockIdx.y];
    const int T = (n + 15) / 16;
    const int npad = T * 16;
    zd* As = reinterpret_cast<zd*>(smem_raw);  // [npad][LDA]
    zd* Bs = As + (size_t)npad * LDA;          // [npad][LDA]
    const bool same = (pr.A == pr.B) && pr.w == nullptr;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
    const size_t k_begin = (size_t)blockIdx.x * kper;
    const size_t k_end = std::min(Ktot, k_begin + kper);

    d4 acc[TILES_PER_WAVE][4];
#pragma unroll
    for (int t = 0; t < TILES_PER_WAVE; t++)
#pragma unroll
        for (int q = 0; q < 4; q++) acc[t][q] = d4{0.0, 0.0, 0.0, 0.0};

    const int ntiles = T * T;
    const int tile0 = blockIdx.z * nwaves * TILES_PER_WAVE;
    for (size_t k0 = k_begin; k0 < k_end; k0 += KC) {
        // stage A and (weighted) B: row i = mode, column kk
        for (int e = threadIdx.x; e < npad * KC; e += blockDim.x) {
            const int i = e / KC, kk = e % KC;
            const size_t k = k0 + kk;
            zd a = cmk<double>(0.0, 0.0), b = cmk<double>(0.0, 0.0);
            if (i < n && k < k_end) {
                a = pr.A[(size_t)i * Ktot + k];
                if (!same) {
                    b = pr.B[(size_t)i * Ktot + k];
                    if (pr.w) b = cmul(pr.w[k % Kw], b);
                }
            }
            As[i * LDA + kk] = a;
            if (!same) Bs[i * LDA + kk] = b;
        }
        __syncthreads();
        const zd* Bsrc = same ? As : Bs;
        for (int t = 0; t < TILES_PER_WAVE; t++) {
            const int tile = tile0 + wave + t * nwaves;
            if (tile >= ntiles) break;
            const int ti = tile / T, tj = tile % T;
#pragma unroll
            for (int kk = 0; kk < KC; kk += 4) {
                // A operand 16x4: lane -> (row i = lane%16, k = lane/16); B operand 4x16: (k = lane/16, col = lane%16)
                const zd a = As[(ti * 16 + (lane & 15)) * LDA + kk + (lane >> 4)];
                const zd b = Bsrc[(tj * 16 + (lane & 15)) * LDA + kk + (lane >> 4)];
                // conj(A)^T B:  re = ar*br + ai*bi,  im = ar*bi - ai*br
                acc[t][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, b.re, acc[t][0], 0, 0, 0);
                acc[t][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.im, b.im, acc[t][1], 0, 0, 0);
                acc[t][2] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, b.im, acc[t][2], 0, 0, 0);
                acc[t][3] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.im, b.re, acc[t][3], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // D layout (v_mfma_f64_16x16x4f64): register v of lane holds D[4*v + lane/16][lane%16]
    zd* out = partial + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * (size_t)n * n;
    for (int t = 0; t < TILES_PER_WAVE; t++) {
        const int tile = tile0 + wave + t * nwaves;
        if (tile >= ntiles) break;
        const int ti = tile / T, tj = tile % T;
        const int j = tj * 16 + (lane & 15);
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int i = ti * 16 + 4 * v + (lane >> 4);
            if (i < n && j < n)
                out[i + (size_t)n * j] = cmk<double>(acc[t][0][v] + acc[t][1][v], acc[t][2][v] - acc[t][3][v]);
        }
    }
}

// Grouped variant: the problems of a group share A and B (every flat air region
// of every frequency shares FFT(Phi) and FFT(mask Phi); every disk region shares
// Phi), so a block stages one chunk of A and B for up to WG_G problems and
// applies each problem's weight to B on the fly.  A wave owns WG_TPW output
// tiles (problem, ti, tj).  Same output layout as k_wgram.
constexpr int WG_G = 4;
constexpr int WG_TPW = 3;

struct GemmGroup {
    const zd* A;
    const zd* B;
    int first;  // index of the group's first problem in the problem list
    int count;  // problems in the group (<= WG_G)
};

constexpr int WG_WAVES = 12;  // waves per block at most (tiles per block <= WG_WAVES * WG_TPW)

// launch shape for T x T output tiles per problem
template <int T>
struct WgShape {
    static constexpr int G = (WG_G < WG_WAVES * WG_TPW / (T * T)) ? WG_G : WG_WAVES * WG_TPW / (T * T);
    static constexpr int WAVES = (G * T * T + WG_TPW - 1) / WG_TPW;
    static constexpr int Q = (16 * T * KC + 64 * WAVES - 1) / (64 * WAVES);  // staged elements per thread
    static_assert(G >= 1 && WAVES <= WG_WAVES && G * KC <= 64 * WAVES, "Gram block shape");
};

// Double-buffered: while the waves run the MFMAs of chunk c from one LDS
// buffer, chunk c+1 travels from HBM into registers; it is written to the
// other buffer afterwards, one barrier per chunk.
template <int T>
__global__ void __launch_bounds__(64 * WgShape<T>::WAVES) k_wgram_grouped(const GemmGroup* __restrict__ groups,
                                                                           const GemmProb* __restrict__ probs, int n,
                                                                           size_t Ktot, size_t Kw, size_t kper,
                                                                           zd* partial) {
    extern __shared__ unsigned char smem_raw[];
    constexpr int WG_Q = WgShape<T>::Q;
    const GemmGroup gr = groups[blockIdx.y];
    constexpr int npad = T * 16, tpp = T * T;
    const size_t buf = (size_t)npad * LDA * 2 + (size_t)WG_G * KC;  // one stage: A, B, weights
    zd* base = reinterpret_cast<zd*>(smem_raw);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
    const size_t k_begin = (size_t)blockIdx.x * kper;
    const size_t k_end = std::min(Ktot, k_begin + kper);
    const int ntiles = gr.count * tpp;
    // the weight vector this thread stages (threads < WG_G * KC)
    const int wg = threadIdx.x / KC;
    const zd* wsrc = (wg < gr.count) ? probs[gr.first + wg].w : nullptr;

    zd ra[WG_Q], rb[WG_Q], rw;
    auto fetch = [&](size_t k0) {
#pragma unroll
        for (int q = 0; q < WG_Q; q++) {
            const int e = threadIdx.x + q * blockDim.x;
            const int i = e / KC, kk = e % KC;
            const size_t k = k0 + kk;
            ra[q] = cmk<double>(0.0, 0.0);
            rb[q] = cmk<double>(0.0, 0.0);
            if (e < npad * KC && i < n && k < k_end) {
                ra[q] = gr.A[(size_t)i * Ktot + k];
                rb[q] = gr.B[(size_t)i * Ktot + k];
            }
        }
        rw = cmk<double>(0.0, 0.0);
        if ((int)threadIdx.x < WG_G * KC) {
            const size_t k = k0 + threadIdx.x % KC;
            if (wg < gr.count && k < k_end) rw = wsrc ? wsrc[k % Kw] : cmk<double>(1.0, 0.0);
        }
    };
    auto store = [&](zd* st) {
        zd* As = st;
        zd* Bs = st + (size_t)npad * LDA;
        zd* Ws = Bs + (size_t)npad * LDA;
#pragma unroll
        for (int q = 0; q < WG_Q; q++) {
            const int e = threadIdx.x + q * blockDim.x;
            if (e < npad * KC) {
                const int i = e / KC, kk = e % KC;
                As[i * LDA + kk] = ra[q];
                Bs[i * LDA + kk] = rb[q];
            }
        }
        if ((int)threadIdx.x < WG_G * KC) Ws[threadIdx.x] = rw;
    };

    d4 acc[WG_TPW][4];
#pragma unroll
    for (int t = 0; t < WG_TPW; t++)
#pragma unroll
        for (int q = 0; q < 4; q++) acc[t][q] = d4{0.0, 0.0, 0.0, 0.0};

    int stage = 0;
    if (k_begin < k_end) {
        fetch(k_begin);
        store(base);
    }
    __syncthreads();
    for (size_t k0 = k_begin; k0 < k_end; k0 += KC) {
        const bool more = k0 + KC < k_end;
        if (more) fetch(k0 + KC);
        const zd* As = base + (size_t)stage * buf;
        const zd* Bs = As + (size_t)npad * LDA;
        const zd* Ws = Bs + (size_t)npad * LDA;
#pragma unroll
        for (int t = 0; t < WG_TPW; t++) {
            const int tile = wave + t * nwaves;
            if (tile >= ntiles) break;
            const int g = tile / tpp, tt = tile % tpp;
            const int ti = tt / T, tj = tt % T;
#pragma unroll
            for (int kk = 0; kk < KC; kk += 4) {
                const int kq = kk + (lane >> 4);
                const zd a = As[(ti * 16 + (lane & 15)) * LDA + kq];
                const zd b = cmul(Ws[g * KC + kq], Bs[(tj * 16 + (lane & 15)) * LDA + kq]);
                acc[t][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, b.re, acc[t][0], 0, 0, 0);
                acc[t][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.im, b.im, acc[t][1], 0, 0, 0);
                acc[t][2] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, b.im, acc[t][2], 0, 0, 0);
                acc[t][3] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.im, b.re, acc[t][3], 0, 0, 0);
            }
        }
        if (more) store(base + (size_t)(stage ^ 1) * buf);
        __syncthreads();  // the next chunk is staged; every wave is done with this one
        stage ^= 1;
    }
    for (int t = 0; t < WG_TPW; t++) {
        const int tile = wave + t * nwaves;
        if (tile >= ntiles) break;
        const int g = tile / tpp, tt = tile % tpp;
        const int ti = tt / T, tj = tt % T;
        zd* out = partial + ((size_t)(gr.first + g) * gridDim.x + blockIdx.x) * (size_t)n * n;
        const int j = tj * 16 + (lane & 15);
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int i = ti * 16 + 4 * v + (lane >> 4);
            if (i < n && j < n)
                out[i + (size_t)n * j] = cmk<double>(acc[t][0][v] + acc[t][1][v], acc[t][2][v] - acc[t][3][v]);
        }
    }
}

// P[out] = sum over splits of partial[prob][split]
__global__ void k_wgram_reduce(const GemmProb* __restrict__ probs, const zd* __restrict__ partial, int splits, int nn,
                               zd* P) {
    const GemmProb pr = probs[blockIdx.y];
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nn; e += gridDim.x * blockDim.x) {
        zd s = cmk<double>(0.0, 0.0);
        for (int q = 0; q < splits; q++) s = cadd(s, partial[((size_t)blockIdx.y * splits + q) * nn + e]);
        P[(size_t)pr.out * nn + e] = s;
    }
}

// ---------------------------------------------------------------------------
// small dense FP64 linear algebra, one block per matrix
// ---------------------------------------------------------------------------
// In-place LU with partial pivoting of an n x n column-major matrix M, then
// solve M X = RHS for nrhs right-hand sides (RHS may be the identity when
// `identity` is set).  Result X written to X (n x nrhs, column-major).
// Visits e = idx, idx + stride, ... < rows * cols as (r, c) with e = r * cols + c, stepping the pair
// instead of dividing every index.
template <typename F>
__device__ __forceinline__ void each2d(int idx, int stride, int rows, int cols, F&& f) {
    int r = idx / cols, c = idx - r * cols;
    const int dr = stride / cols, dc = stride - dr * cols;
    while (r < rows) {
        f(r, c);
        c += dc;
        r += dr;
        if (c >= cols) {
            c -= cols;
            r++;
        }
    }
}
template <typename F>
__device__ __forceinline__ void wave_each2d(int lane, int rows, int cols, F&& f) {
    each2d(lane, 64, rows, cols, f);
}

__device__ void lu_factor_block(zd* M, int n, int* piv, double* sval, int* sidx) {
    for (int k = 0; k < n; k++) {
        // pivot: first index of max |re|+|im| in column k, rows k..n-1
        double best = -1.0;
        int bi = k;
        for (int i = k + threadIdx.x; i < n; i += blockDim.x) {
            const zd v = M[i + (size_t)n * k];
            const double a = fabs(v.re) + fabs(v.im);
            if (a > best) { best = a; bi = i; }
        }
        sval[threadIdx.x] = best;
        sidx[threadIdx.x] = bi;
        __syncthreads();
        if (threadIdx.x == 0) {  // only the first n - k threads hold candidates
            double b = -1.0;
            int p = k;
            const int nt = min((int)blockDim.x, n - k);
            for (int t = 0; t < nt; t++)
                if (sval[t] > b || (sval[t] == b && sidx[t] < p)) { b = sval[t]; p = sidx[t]; }
            piv[k] = p;
        }
        __syncthreads();
        const int p = piv[k];
        if (p != k)
            for (int j = threadIdx.x; j < n; j += blockDim.x) {
                const zd t = M[k + (size_t)n * j];
                M[k + (size_t)n * j] = M[p + (size_t)n * j];
                M[p + (size_t)n * j] = t;
            }
        __syncthreads();
        const zd d = M[k + (size_t)n * k];
        for (int i = k + 1 + threadIdx.x; i < n; i += blockDim.x) {
            const zd v = M[i + (size_t)n * k];
            // v / d
            const double den = d.re * d.re + d.im * d.im;
            M[i + (size_t)n * k] = cmk<double>((v.re * d.re + v.im * d.im) / den, (v.im * d.re - v.re * d.im) / den);
        }
        __syncthreads();
        const int m = n - k - 1;
        if (m > 0)
            each2d(threadIdx.x, blockDim.x, m, m, [&](int jj, int ii) {
                const int i = k + 1 + ii, j = k + 1 + jj;
                M[i + (size_t)n * j] = csub(M[i + (size_t)n * j], cmul(M[i + (size_t)n * k], M[k + (size_t)n * j]));
            });
        __syncthreads();
    }
}

// forward/back substitution for column c (one thread per right-hand side)
__device__ void lu_solve_col(const zd* M, int n, const int* piv, zd* x) {
    for (int k = 0; k < n; k++)
        if (piv[k] != k) { const zd t = x[k]; x[k] = x[piv[k]]; x[piv[k]] = t; }
    for (int i = 0; i < n; i++)
        for (int k = 0; k < i; k++) x[i] = csub(x[i], cmul(M[i + (size_t)n * k], x[k]));
    for (int i = n - 1; i >= 0; i--) {
        for (int k = i + 1; k < n; k++) x[i] = csub(x[i], cmul(M[i + (size_t)n * k], x[k]));
        const zd d = M[i + (size_t)n * i];
        const double den = d.re * d.re + d.im * d.im;
        x[i] = cmk<double>((x[i].re * d.re + x[i].im * d.im) / den, (x[i].im * d.re - x[i].re * d.im) / den);
    }
}

// inverses of a batch of n x n matrices: blockIdx.x = matrix.  work: [batch][n][n] + [batch][n][n] (columns)
__global__ void k_batched_inverse(const zd* __restrict__ P, zd* Pinv, int n, zd* work, const int* __restrict__ sel) {
    __shared__ double sval[256];
    __shared__ int sidx[256];
    __shared__ int piv[BF_MAX_MODES];
    const int b = sel ? sel[blockIdx.x] : blockIdx.x;
    const size_t nn = (size_t)n * n;
    zd* M = work + (size_t)blockIdx.x * nn;
    for (size_t e = threadIdx.x; e < nn; e += blockDim.x) M[e] = P[(size_t)b * nn + e];
    __syncthreads();
    lu_factor_block(M, n, piv, sval, sidx);
    zd* X = Pinv + (size_t)b * nn;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        zd* x = X + (size_t)n * j;
        for (int i = 0; i < n; i++) x[i] = cmk<double>(i == j ? 1.0 : 0.0, 0.0);
        lu_solve_col(M, n, piv, x);
    }
}

// The same inverses with one wave per matrix and the matrix and the solutions in
// LDS (n <= INV_WAVE_MAX): the pivot is a wave argmax (largest |re|+|im|, lowest
// row on ties, as above), every arithmetic step is the one of lu_factor_block /
// lu_solve_col in the same order, so the results are identical.
constexpr int INV_WAVE_MAX = 48;

// Matrix index of block (x, y): sel[x], or, with the (base, stride) list `bs`, bs[x].base + y bs[x].stride
// (a region's matrix at frequency y), or x.
struct InvSel {
    int64_t base;
    int64_t stride;
};
// Gauss-Jordan inversion in place with partial (row) pivoting, one wave per
// matrix held in LDS (n^2 complex): for each k the pivot row (first row of
// largest |re| + |im| in column k, as izamax) is swapped in, scaled by the
// reciprocal of the pivot, and column k is eliminated from every other row;
// the row swaps are undone on the columns when the result is written.  n
// steps of three barriers each, every lane busy in the eliminations.
__global__ void __launch_bounds__(64) k_batched_inverse_wave(const zd* __restrict__ P, zd* Pinv, int n,
                                                           const int* __restrict__ sel,
                                                           const InvSel* __restrict__ bs) {
    extern __shared__ unsigned char smem_raw[];
    zd* M = reinterpret_cast<zd*>(smem_raw);  // column-major n x n
    __shared__ zd colk[INV_WAVE_MAX];
    __shared__ int piv[INV_WAVE_MAX], sig[INV_WAVE_MAX];
    const int lane = threadIdx.x;
    const int64_t b = bs ? bs[blockIdx.x].base + (int64_t)blockIdx.y * bs[blockIdx.x].stride
                         : (sel ? sel[blockIdx.x] : blockIdx.x);
    const int nn = n * n;
    const zd* src = P + (size_t)b * nn;
    for (int e = lane; e < nn; e += 64) M[e] = src[e];
    __syncthreads();
    for (int k = 0; k < n; k++) {
        double best = -1.0;
        int bi = k;
        for (int i = k + lane; i < n; i += 64) {
            const zd v = M[i + n * k];
            const double a = fabs(v.re) + fabs(v.im);
            if (a > best) { best = a; bi = i; }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const double ob = __shfl_xor(best, o);
            const int oi = __shfl_xor(bi, o);
            if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
        }
        const int p = bi;
        if (p != k)
            for (int j = lane; j < n; j += 64) {
                const zd t = M[k + n * j];
                M[k + n * j] = M[p + n * j];
                M[p + n * j] = t;
            }
        if (lane == 0) piv[k] = p;
        __syncthreads();
        // r = 1 / pivot; row k scaled by r (its pivot entry becomes r); column k kept for the elimination
        const zd d = M[k + n * k];
        const double den = d.re * d.re + d.im * d.im;
        const zd r = cmk<double>(d.re / den, -d.im / den);
        for (int i = lane; i < n; i += 64)
            if (i != k) colk[i] = M[i + n * k];
        for (int j = lane; j < n; j += 64) M[k + n * j] = j == k ? r : cmul(M[k + n * j], r);
        __syncthreads();
        // row i != k: M[i][j] -= M_old[i][k] M[k][j] (j != k), M[i][k] = -M_old[i][k] r
        wave_each2d(lane, n, n - 1, [&](int j, int ii) {
            const int i = ii < k ? ii : ii + 1;
            const zd f = colk[i];
            M[i + n * j] = j == k ? csub(cmk<double>(0.0, 0.0), cmul(f, r)) : csub(M[i + n * j], cmul(f, M[k + n * j]));
        });
        __syncthreads();
    }
    // undo the row swaps on the columns, in reverse order: result column j is column sig[j]
    if (lane == 0) {
        for (int j = 0; j < n; j++) sig[j] = j;
        for (int k = n - 1; k >= 0; k--) {
            const int t = sig[k];
            sig[k] = sig[piv[k]];
            sig[piv[k]] = t;
        }
    }
    __syncthreads();
    zd* dst = Pinv + (size_t)b * nn;
    wave_each2d(lane, n, n, [&](int j, int i) { dst[i + n * j] = M[i + n * sig[j]]; });
}

// ---------------------------------------------------------------------------
// transfer-matrix chain, one block per job (src/boost_fractor_transformer.jl:326-357)
// ---------------------------------------------------------------------------
__global__ void k_chain(const ChainJob* __restrict__ jobs, const ChainStep* __restrict__ steps,
                        const zd* __restrict__ P, const zd* __restrict__ Pinv, int n, const zd* __restrict__ e0_all,
                        const zd* __restrict__ reflect, zd* Tbuf, zd* work, zd* boost, zd* refl) {
    __shared__ double sval[256];
    __shared__ int sidx[256];
    __shared__ int piv[BF_MAX_MODES];
    const int f = blockIdx.x;
    const ChainJob job = jobs[f];
    const int nsteps = job.nsteps;
    const zd* e0 = e0_all + job.e0off;
    const int N2 = 2 * n;
    const size_t nn = (size_t)n * n, NN = (size_t)N2 * N2;
    zd* T = Tbuf + (size_t)f * 2 * NN;
    zd* Tn = T + NN;
    zd* ab = work + (size_t)f * (nn + 3 * (size_t)n);  // [n]
    zd* M = ab + n;                                     // [n][n]
    zd* tmp = M + nn;                                   // [2n]
    for (size_t e = threadIdx.x; e < NN; e += blockDim.x) T[e] = cmk<double>((e % N2) == (e / N2) ? 1.0 : 0.0, 0.0);
    for (int i = threadIdx.x; i < n; i += blockDim.x) ab[i] = cmk<double>(0.0, 0.0);
    __syncthreads();
    for (int q = 0; q < nsteps; q++) {
        const ChainStep st = steps[job.step0 + q];
        // axion_beam += S * (T21*e0 + T22*e0)
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            zd s21 = cmk<double>(0.0, 0.0), s22 = cmk<double>(0.0, 0.0);
            for (int k = 0; k < n; k++) {
                s21 = cadd(s21, cmul(T[(n + i) + (size_t)N2 * k], e0[k]));
                s22 = cadd(s22, cmul(T[(n + i) + (size_t)N2 * (n + k)], e0[k]));
            }
            ab[i] = cadd(ab[i], cmul(st.S, cadd(s21, s22)));
        }
        // T <- T * [a P, b Pinv; b P, a Pinv]
        const zd* Pm = P + (size_t)(job.pbase + st.reg) * nn;
        const zd* Pi = Pinv + (size_t)(job.pbase + st.reg) * nn;
        for (size_t e = threadIdx.x; e < NN; e += blockDim.x) {
            const int i = (int)(e % N2), j = (int)(e / N2);
            const bool left = j < n;
            const int jj = left ? j : j - n;
            const zd ca = left ? st.a : st.b, cb = left ? st.b : st.a;
            const zd* Mb = left ? Pm : Pi;
            zd s = cmk<double>(0.0, 0.0);
            for (int k = 0; k < n; k++) {
                const zd t = cadd(cmul(ca, T[i + (size_t)N2 * k]), cmul(cb, T[i + (size_t)N2 * (n + k)]));
                s = cadd(s, cmul(t, Mb[k + (size_t)n * jj]));
            }
            Tn[e] = s;
        }
        __syncthreads();
        for (size_t e = threadIdx.x; e < NN; e += blockDim.x) T[e] = Tn[e];
        __syncthreads();
    }
    // boost = -T22 \ axion_beam ;  refl = -T22 \ (T21 * reflect)
    for (size_t e = threadIdx.x; e < nn; e += blockDim.x) {
        const int i = (int)(e % n), j = (int)(e / n);
        M[e] = T[(n + i) + (size_t)N2 * (n + j)];
    }
    __syncthreads();
    lu_factor_block(M, n, piv, sval, sidx);
    if (threadIdx.x == 0) {
        for (int i = 0; i < n; i++) tmp[i] = ab[i];
        lu_solve_col(M, n, piv, tmp);
        for (int i = 0; i < n; i++) boost[(size_t)job.out + i] = cmk<double>(-tmp[i].re, -tmp[i].im);
        if (reflect) {
            for (int i = 0; i < n; i++) {
                zd s = cmk<double>(0.0, 0.0);
                for (int k = 0; k < n; k++) s = cadd(s, cmul(T[(n + i) + (size_t)N2 * k], reflect[k]));
                tmp[n + i] = s;
            }
            lu_solve_col(M, n, piv, tmp + n);
            for (int i = 0; i < n; i++) refl[(size_t)job.out + i] = cmk<double>(-tmp[n + i].re, -tmp[n + i].im);
        }
    }
}

// Entry g = i + n j of the matrix (inv = false) or of the inverse of matrix slot m of a chain step; a
// disk step forms it from G, Ginv, the slot's phases and their reciprocals
__device__ __forceinline__ zd chain_entry(const ChainSrc& S, int disk, int64_t m, int n, int g, int i, int j,
                                         bool inv) {
    if (!disk) return inv ? S.Pinv[m * n * n + g] : S.P[m * n * n + g];
    if (!inv) return cmul(S.ph[m * n + i], S.G[g]);
    return cmul(S.Ginv[g], S.phinv[m * n + j]);
}

// The chain in LDS for n <= CHAIN_LDS_MAX.  Only the bottom block row [T21 T22]
// of T feeds the axion sum and the solves, and it updates by itself:
//     T21' = U P,  T22' = V inv(P),  U = a T21 + b T22,  V = b T21 + a T22
// (16 n^3 flop per region instead of 64 n^3 for the whole 2n x 2n product).
// LDS holds U and V of the coming step, each entry formed once by the thread
// that produced it from the previous step's products (T21 and T22 themselves
// after the last step).  Since a + b = 1 (a = (n1 + n0) / (2 n1),
// b = (n1 - n0) / (2 n1)), U + V = T21 + T22 and the axion sum
// S (T21 + T22) e0 is S (U e0 + V e0): e0 sits in LDS as column n of both P
// and inv(P), and whenever the thread grid has a spare column
// (floor(256 / n) ceil(n / floor(256 / n)) > n: n = 9, 15, 25, ...) the
// products form it on the way; otherwise the row's first column group adds it.
// Every output sums over k in increasing order (fused multiply-adds).  Thread
// t owns row t % n and the columns t / n + c floor(256 / n) of both products
// (all 256 threads busy for n = 16, 25, 32, ...), and loads the next region's
// P and inv(P) into registers while it computes, so each region costs two
// barriers and no exposed HBM latency.  (Loading each k's operands one k ahead
// as well measured slower: 230 instead of 140 VGPRs.)  LDS: U, V (n^2 complex
// each), P, inv(P) (n (n + 1) complex each, column n = e0).  NM: the largest n
// of the instantiation (its register arrays).
constexpr int CHAIN_LDS_MAX = 48;

// column groups and columns per thread of k_chain_lds for n modes
__host__ __device__ constexpr int chain_groups(int n) { return 256 / n; }
__host__ __device__ constexpr int chain_cols(int n) { return (n + chain_groups(n) - 1) / chain_groups(n); }
// the largest n <= NM with chain_cols(n) == NC
__host__ __device__ constexpr int chain_nmax(int NM, int NC) {
    int m = 0;
    for (int n = 1; n <= NM; n++)
        if (chain_cols(n) == NC) m = n;
    return m;
}

template <int NM, int NC>
struct ChainShape {
    static constexpr bool PF = NM <= 32;  // next region's matrices through registers (else loaded after the barrier)
    static constexpr int NX = chain_nmax(NM, NC);
    static constexpr int PPT = PF ? (2 * NX * NX + 255) / 256 : 1;  // P and inv(P) entries per thread
};

// columns of P and inv(P) in the LDS of k_chain_lds: n, e0, and padding up to the thread grid's
// floor(256 / n) ceil(n / floor(256 / n)) columns (read, never used)
__host__ __device__ constexpr int chain_width(int n) {
    return n + 1 > chain_groups(n) * chain_cols(n) ? n + 1 : chain_groups(n) * chain_cols(n);
}
// dynamic LDS of k_chain_lds: U, V, P, inv(P), then the solves' vectors and pivot scratch
__host__ __device__ inline size_t chain_lds_bytes(int n) {
    return sizeof(zd) * ((size_t)2 * n * n + (size_t)2 * n * chain_width(n) + 3 * n) +
           (sizeof(double) + sizeof(int)) * 256 + sizeof(int) * n;
}

template <int NM, int NC>
__global__ void __launch_bounds__(256) k_chain_lds(const ChainJob* __restrict__ jobs,
                                                 const ChainStep* __restrict__ steps, const ChainSrc src, int n,
                                                 const zd* __restrict__ e0_all, const zd* __restrict__ reflect,
                                                 zd* boost, zd* refl) {
    using CS = ChainShape<NM, NC>;
    extern __shared__ unsigned char smem_raw[];
    const int nn = n * n;
    zd* U = reinterpret_cast<zd*>(smem_raw);  // column-major n x n: [i + n k]; T21 after the last step
    zd* V = U + nn;                           // T22 after the last step
    const int nw = n * chain_width(n);
    zd* Pl = V + nn;                          // P of the current region, n x chain_width(n), column n = e0
    zd* Pil = Pl + nw;                        // inv(P), likewise
    zd* ab = Pil + nw;                        // axion_beam
    zd* tmp = ab + n;                         // right-hand sides of the solves
    double* sval = reinterpret_cast<double*>(tmp + 2 * n);
    int* sidx = reinterpret_cast<int*>(sval + 256);
    int* piv = sidx + 256;
    const ChainJob job = jobs[blockIdx.x];
    const int tid = threadIdx.x;
    const int ng = 256 / n, row = tid % n, cg = tid / n;  // NC = ceil(n / ng)
    const bool act = cg < ng;
    const bool spare = ng * NC > n;                 // block-uniform
    const int own = spare ? n - ng * (NC - 1) : 0;  // column group of the row's axion sum
    {
        // T21 = 0, T22 = I: U = b I, V = a I for the first step
        const zd z = cmk<double>(0.0, 0.0);
        zd du = z, dv = cmk<double>(1.0, 0.0);
        if (job.nsteps > 0) {
            du = steps[job.step0].b;
            dv = steps[job.step0].a;
        }
        for (int e = tid; e < nn; e += 256) {
            const bool d = (e % n) == (e / n);
            U[e] = cmk<double>(d ? du.re : 0.0, d ? du.im : 0.0);
            V[e] = cmk<double>(d ? dv.re : 0.0, d ? dv.im : 0.0);
        }
    }
    for (int i = tid; i < n; i += 256) Pl[nn + i] = Pil[nn + i] = e0_all[job.e0off + i];
    // entry r < 2 nn of [P; inv(P)] of step q
    auto entry = [&](const ChainStep& sq, int r) -> zd {
        const bool inv = r >= nn;
        const int g = inv ? r - nn : r;
        return chain_entry(src, sq.disk, job.pbase + sq.reg, n, g, g % n, g / n, inv);
    };
    auto lds_at = [&](int r) -> int { return r < nn ? r : r + nw - nn; };  // entry r of [P; inv(P)] in Pl
    if (job.nsteps > 0) {
        const ChainStep s0 = steps[job.step0];
        for (int r = tid; r < 2 * nn; r += 256) Pl[lds_at(r)] = entry(s0, r);
    }
    zd abr = cmk<double>(0.0, 0.0);  // axion_beam[row] in the thread of column group `own`
    __syncthreads();
    for (int q = 0; q < job.nsteps; q++) {
        const ChainStep st = steps[job.step0 + q];
        const bool more = q + 1 < job.nsteps;
        const ChainStep sn = more ? steps[job.step0 + q + 1] : st;
        zd pn[CS::PPT];
        if constexpr (CS::PF) {
            // every element assigned on every path, so the array stays in registers
#pragma unroll
            for (int i = 0; i < CS::PPT; i++) {
                const int r = tid + 256 * i;
                pn[i] = more && r < 2 * nn ? entry(sn, r) : cmk<double>(0.0, 0.0);
            }
        }
        zd o21[NC], o22[NC];
#pragma unroll
        for (int c = 0; c < NC; c++) o21[c] = o22[c] = cmk<double>(0.0, 0.0);
        if (act) {
            // columns beyond n read column n (their sums are never stored): no divergent branches
            // thread (row, cg) reads U[row, k], V[row, k] and column cg + ng c of P and inv(P): the
            // columns beyond n (padding) give sums that are never stored
            const zd* Ur = U + row;
            const zd* Vr = V + row;
            const zd* Pc = Pl + cg * n;
            const int cs = ng * n;  // LDS distance between a thread's columns
            auto kstep = [&](int k) {
                const zd u = Ur[n * k], v = Vr[n * k];
                zd p[NC], pi[NC];
#pragma unroll
                for (int c = 0; c < NC; c++) {
                    p[c] = Pc[k + cs * c];
                    pi[c] = Pc[nw + k + cs * c];
                }
#pragma unroll
                for (int c = 0; c < NC; c++) {
                    o21[c] = cmac(o21[c], u, p[c]);
                    o22[c] = cmac(o22[c], v, pi[c]);
                }
            };
#pragma unroll 2
            for (int k = 0; k < n; k++) kstep(k);
            if (spare) {
                if (cg == own) abr = cadd(abr, cmul(st.S, cadd(o21[NC - 1], o22[NC - 1])));
            } else if (cg == 0) {
                zd s = cmk<double>(0.0, 0.0);
                for (int k = 0; k < n; k++) s = cmac(s, cadd(U[row + n * k], V[row + n * k]), Pl[nn + k]);
                abr = cadd(abr, cmul(st.S, s));
            }
        }
        __syncthreads();
        if (act) {
#pragma unroll
            for (int c = 0; c < NC; c++) {
                const int j = cg + ng * c;
                if (j < n) {
                    zd t21 = o21[c], t22 = o22[c];
                    if (more) {
                        t21 = cadd(cmul(sn.a, o21[c]), cmul(sn.b, o22[c]));
                        t22 = cadd(cmul(sn.b, o21[c]), cmul(sn.a, o22[c]));
                    }
                    U[row + n * j] = t21;
                    V[row + n * j] = t22;
                }
            }
        }
        if constexpr (CS::PF) {
#pragma unroll
            for (int i = 0; i < CS::PPT; i++) {
                const int r = tid + 256 * i;
                if (more && r < 2 * nn) Pl[lds_at(r)] = pn[i];
            }
        } else if (more) {
            for (int r = tid; r < 2 * nn; r += 256) Pl[lds_at(r)] = entry(sn, r);
        }
        __syncthreads();
    }
    if (act && cg == own) ab[row] = abr;
    zd* T21 = U;
    zd* T22 = V;
    // boost = -T22 \ axion_beam ;  refl = -T22 \ (T21 * reflect)
    if (reflect)
        for (int i = tid; i < n; i += 256) {
            zd s = cmk<double>(0.0, 0.0);
            for (int k = 0; k < n; k++) s = cadd(s, cmul(T21[i + n * k], reflect[k]));
            tmp[n + i] = s;
        }
    __syncthreads();
    lu_factor_block(T22, n, piv, sval, sidx);
    if (tid == 0) {
        for (int i = 0; i < n; i++) tmp[i] = ab[i];
        lu_solve_col(T22, n, piv, tmp);
        for (int i = 0; i < n; i++) boost[(size_t)job.out + i] = cmk<double>(-tmp[i].re, -tmp[i].im);
        if (reflect) {
            lu_solve_col(T22, n, piv, tmp + n);
            for (int i = 0; i < n; i++) refl[(size_t)job.out + i] = cmk<double>(-tmp[n + i].re, -tmp[n + i].im);
        }
    }
}

// The chain for the mode counts of the production sweeps (9, 15, 25, 45), register tiled.  k_chain_lds
// is bound by LDS issue: per k a thread reads U and V of its row and C entries each of P and inv(P) for
// 2 C complex multiply-adds.  Here thread t of NT owns rows rg + NRG r (r < R) and columns cg + NCG c
// (c < C), rg = t % NRG, cg = t / NRG, of both products, and per k reads R entries each of U and V and
// C each of P and inv(P) for 2 R C multiply-adds.  The columns run to NF (column NF of P and inv(P) is
// e0: the axion sum); rows and columns past that are padding whose sums are never stored.  Layout,
// steps and the order of every sum are those of k_chain_lds.
template <int NF, int R, int C, int NT>
struct ChainTile {
    static constexpr int NRG = (NF + R - 1) / R;      // row groups
    static constexpr int NCG = (NF + 1 + C - 1) / C;  // column groups over columns 0 .. NF
    static constexpr int W = NCG * C;                 // LDS columns of P and inv(P)
    static constexpr int ACT = NRG * NCG;             // threads with a tile
    static constexpr int OWN_CG = NF % NCG, OWN_C = NF / NCG;  // tile holding column NF
    static constexpr int PPT = (2 * NF * NF + NT - 1) / NT;    // next region's entries per thread
    static constexpr int UK = NF % 5 == 0 ? 5 : (NF % 3 == 0 ? 3 : 1);  // k unroll
    static_assert(ACT <= NT, "tile grid exceeds the block");
    static constexpr size_t lds_bytes() {
        return sizeof(zd) * ((size_t)2 * NF * NF + (size_t)2 * NF * W + 3 * NF) + (sizeof(double) + sizeof(int)) * NT +
               sizeof(int) * NF;
    }
};

template <int NF, int R, int C, int NT>
__global__ void __launch_bounds__(NT) k_chain_tile(const ChainJob* __restrict__ jobs,
                                                   const ChainStep* __restrict__ steps, const ChainSrc src,
                                                   const zd* __restrict__ e0_all, const zd* __restrict__ reflect,
                                                   zd* boost, zd* refl) {
    using T = ChainTile<NF, R, C, NT>;
    constexpr int n = NF, nn = NF * NF, nw = NF * T::W;
    extern __shared__ unsigned char smem_raw[];
    zd* U = reinterpret_cast<zd*>(smem_raw);  // column-major n x n; T21 after the last step
    zd* V = U + nn;                           // T22 after the last step
    zd* Pl = V + nn;                          // P of the current region, n x W, column n = e0
    zd* Pil = Pl + nw;                        // inv(P), likewise
    zd* ab = Pil + nw;
    zd* tmp = ab + n;
    double* sval = reinterpret_cast<double*>(tmp + 2 * n);
    int* sidx = reinterpret_cast<int*>(sval + NT);
    int* piv = sidx + NT;
    const ChainJob job = jobs[blockIdx.x];
    const int tid = threadIdx.x;
    const int rg = tid % T::NRG, cg = tid / T::NRG;
    const bool act = tid < T::ACT;
    {
        const zd z = cmk<double>(0.0, 0.0);
        zd du = z, dv = cmk<double>(1.0, 0.0);
        if (job.nsteps > 0) {
            du = steps[job.step0].b;
            dv = steps[job.step0].a;
        }
        for (int e = tid; e < nn; e += NT) {
            const bool d = (e % n) == (e / n);
            U[e] = cmk<double>(d ? du.re : 0.0, d ? du.im : 0.0);
            V[e] = cmk<double>(d ? dv.re : 0.0, d ? dv.im : 0.0);
        }
    }
    for (int i = tid; i < n; i += NT) Pl[nn + i] = Pil[nn + i] = e0_all[job.e0off + i];
    auto lds_at = [&](int r) -> int { return r < nn ? r : r + nw - nn; };
    auto entry = [&](const ChainStep& sq, int r) -> zd {  // entry r < 2 nn of [P; inv(P)] of step sq
        const bool inv = r >= nn;
        const int g = inv ? r - nn : r;
        return chain_entry(src, sq.disk, job.pbase + sq.reg, n, g, g % n, g / n, inv);
    };
    if (job.nsteps > 0) {
        const ChainStep s0 = steps[job.step0];
        for (int r = tid; r < 2 * nn; r += NT) Pl[lds_at(r)] = entry(s0, r);
    }
    zd abr[R];
#pragma unroll
    for (int r = 0; r < R; r++) abr[r] = cmk<double>(0.0, 0.0);
    __syncthreads();
    for (int q = 0; q < job.nsteps; q++) {
        const ChainStep st = steps[job.step0 + q];
        const bool more = q + 1 < job.nsteps;
        const ChainStep sn = more ? steps[job.step0 + q + 1] : st;
        zd pn[T::PPT];
#pragma unroll
        for (int i = 0; i < T::PPT; i++) {
            const int r = tid + NT * i;
            pn[i] = more && r < 2 * nn ? entry(sn, r) : cmk<double>(0.0, 0.0);
        }
        zd o21[R][C], o22[R][C];
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int c = 0; c < C; c++) o21[r][c] = o22[r][c] = cmk<double>(0.0, 0.0);
        if (act) {
            const zd* Ur = U + rg;
            const zd* Vr = V + rg;
            const zd* Pc = Pl + cg * n;
#pragma unroll 1
            for (int k0 = 0; k0 < n; k0 += T::UK) {
#pragma unroll
                for (int kk = 0; kk < T::UK; kk++) {
                    const int k = k0 + kk;
                    zd u[R], v[R], p[C], pi[C];
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        u[r] = Ur[T::NRG * r + n * k];
                        v[r] = Vr[T::NRG * r + n * k];
                    }
#pragma unroll
                    for (int c = 0; c < C; c++) {
                        p[c] = Pc[k + T::NCG * n * c];
                        pi[c] = Pc[nw + k + T::NCG * n * c];
                    }
#pragma unroll
                    for (int r = 0; r < R; r++)
#pragma unroll
                        for (int c = 0; c < C; c++) {
                            o21[r][c] = cmac(o21[r][c], u[r], p[c]);
                            o22[r][c] = cmac(o22[r][c], v[r], pi[c]);
                        }
                }
            }
            if (cg == T::OWN_CG) {
#pragma unroll
                for (int r = 0; r < R; r++)
                    abr[r] = cadd(abr[r], cmul(st.S, cadd(o21[r][T::OWN_C], o22[r][T::OWN_C])));
            }
        }
        __syncthreads();
        if (act) {
#pragma unroll
            for (int r = 0; r < R; r++)
#pragma unroll
                for (int c = 0; c < C; c++) {
                    const int i = rg + T::NRG * r, j = cg + T::NCG * c;
                    if (i < n && j < n) {
                        zd t21 = o21[r][c], t22 = o22[r][c];
                        if (more) {
                            t21 = cadd(cmul(sn.a, o21[r][c]), cmul(sn.b, o22[r][c]));
                            t22 = cadd(cmul(sn.b, o21[r][c]), cmul(sn.a, o22[r][c]));
                        }
                        U[i + n * j] = t21;
                        V[i + n * j] = t22;
                    }
                }
        }
#pragma unroll
        for (int i = 0; i < T::PPT; i++) {
            const int r = tid + NT * i;
            if (more && r < 2 * nn) Pl[lds_at(r)] = pn[i];
        }
        __syncthreads();
    }
    if (act && cg == T::OWN_CG) {
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int i = rg + T::NRG * r;
            if (i < n) ab[i] = abr[r];
        }
    }
    zd* T21 = U;
    zd* T22 = V;
    // boost = -T22 \ axion_beam ;  refl = -T22 \ (T21 * reflect)
    if (reflect)
        for (int i = tid; i < n; i += NT) {
            zd s = cmk<double>(0.0, 0.0);
            for (int k = 0; k < n; k++) s = cadd(s, cmul(T21[i + n * k], reflect[k]));
            tmp[n + i] = s;
        }
    __syncthreads();
    lu_factor_block(T22, n, piv, sval, sidx);
    if (tid == 0) {
        for (int i = 0; i < n; i++) tmp[i] = ab[i];
        lu_solve_col(T22, n, piv, tmp);
        for (int i = 0; i < n; i++) boost[(size_t)job.out + i] = cmk<double>(-tmp[i].re, -tmp[i].im);
        if (reflect) {
            lu_solve_col(T22, n, piv, tmp + n);
            for (int i = 0; i < n; i++) refl[(size_t)job.out + i] = cmk<double>(-tmp[n + i].re, -tmp[n + i].im);
        }
    }
}

template <int NF, int R, int C, int NT>
static void launch_chain_tile(hipStream_t s, size_t nj, const ChainJob* jobs, const ChainStep* steps,
                              const ChainSrc& src, const zd* e0, const zd* reflect, zd* boost, zd* refl) {
    const size_t smem = ChainTile<NF, R, C, NT>::lds_bytes();
    allow_lds((const void*)k_chain_tile<NF, R, C, NT>, smem);
    hipLaunchKernelGGL((k_chain_tile<NF, R, C, NT>), dim3((unsigned)nj), dim3(NT), smem, s, jobs, steps, src, e0,
                       reflect, boost, refl);
}

// The chain on the matrix cores for n <= 32 (v_mfma_f64_16x16x4f64).  The operands are zero-padded to
// NP x NP (NP = 16 or 32); wave w of the (NP / 16)^2 owns output tile (w % (NP / 16), w / (NP / 16)) of
// both products and forms it with four MFMAs per complex 4-k step (re += Ar Br + (-Ai) Bi,
// im += Ar Bi + Ai Br) or, GAUSS, three (Ar Br, Ai Bi and (Ar + Ai)(Br + Bi)).  The A operands U and V
// live in LDS, column-major with leading dimension NP + 1 (every operand read is 16 consecutive entries
// per k, conflict-free), in two buffers: region q reads one and writes the next region's U and V into
// the other, so one barrier per region suffices (a buffer is rewritten only after every wave has passed
// the barrier that ends the region reading it).  The B operands, P and inv(P), never enter LDS: each
// lane loads its own entries P[k][j] (k = 4 s + lane / 16, j = the tile's column lane % 16) from global
// memory into registers one region ahead.  Column n of the B operands is e0 when n < NP (the axion sum
// comes out of the products), else the first wave adds it by a dot product.  The padding stays zero:
// U and V are written for rows and columns below n only, and B entries beyond n are zero.  Each MFMA
// sums its four k in an order of its own, so the results agree with the VALU kernels to rounding.
template <int NP>
struct ChainMfma {
    static constexpr int TW = NP / 16;        // tiles per dimension
    static constexpr int NT = 64 * TW * TW;   // one wave per tile
    static constexpr int LD = NP + 1;         // LDS leading dimension (entries)
    static constexpr int KS = NP / 4;         // MFMA k steps
    static size_t lds_bytes(int n) {
        return sizeof(zd) * ((size_t)4 * NP * LD + 3 * (size_t)n) + (sizeof(double) + sizeof(int)) * NT +
               sizeof(int) * (size_t)n;
    }
};

// KF > 0: exactly KF k steps (n in (4 KF - 4, 4 KF]), so the next region's loads are issued in the same
// order on every path and the first use waits for its own load only; KF = 0: k steps from n at run time.
template <int NP, bool GAUSS, int KF>
__global__ void __launch_bounds__(ChainMfma<NP>::NT, 2) k_chain_mfma(const ChainJob* __restrict__ jobs,
                                                                  const ChainStep* __restrict__ steps,
                                                                  const ChainSrc src, int n,
                                                                  const zd* __restrict__ e0_all,
                                                                  const zd* __restrict__ reflect, zd* boost, zd* refl) {
    using T = ChainMfma<NP>;
    constexpr int LD = T::LD, NT = T::NT, TW = T::TW, KS = KF > 0 ? KF : T::KS;
    constexpr int SZ = NP * LD;
    extern __shared__ unsigned char smem_raw[];
    zd* buf = reinterpret_cast<zd*>(smem_raw);  // U of buffer b at buf + 2 b SZ, V at buf + (2 b + 1) SZ
    zd* ab = buf + 4 * SZ;
    zd* tmp = ab + n;
    double* sval = reinterpret_cast<double*>(tmp + 2 * n);
    int* sidx = reinterpret_cast<int*>(sval + NT);
    int* piv = sidx + NT;
    const ChainJob job = jobs[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tr = wave % TW, tc = wave / TW;  // output tile
    const int nn = n * n;
    const bool spare = n < NP;
    const zd* e0 = e0_all + job.e0off;
    {
        zd du = cmk<double>(0.0, 0.0), dv = cmk<double>(1.0, 0.0);
        if (job.nsteps > 0) {
            du = steps[job.step0].b;
            dv = steps[job.step0].a;
        }
        for (int e = tid; e < 4 * SZ; e += NT) {
            const int b = e / SZ, r = e % SZ, i = r % LD, k = r / LD;
            const bool d = b < 2 && i == k && i < n;
            const zd dd = b == 0 ? du : dv;
            buf[e] = cmk<double>(d ? dd.re : 0.0, d ? dd.im : 0.0);
        }
    }
    zd abr[4];
#pragma unroll
    for (int v = 0; v < 4; v++) abr[v] = cmk<double>(0.0, 0.0);
    const bool own = spare && tc == n / 16 && (lane & 15) == n % 16;  // lanes holding column n
    const int ar = tr * 16 + (lane & 15), bc = tc * 16 + (lane & 15), kq = lane >> 4;
    const int kp = (n + 3) & ~3;  // k steps cover k < n
    // The lane's B entries of the coming region: P[k][bc] and inv(P)[k][bc] for k = 4 s + kq, e0[k] in
    // column n.  Indices are clamped into the matrix instead of zeroing: rows k >= n meet the zero columns
    // of U and V, and columns beyond n of the products are never used.  A gap's entry s is a plain load
    // (one address per lane, e0's in column n) issued right after the current region's MFMAs of step s
    // have read the register, and first waited for where the next region uses it; a disk's entries are
    // formed from G, inv(G) and the phases (L2-resident) after the current region's products.
    const int cc = bc < n ? bc : n - 1;
    const bool col = bc == n;
    zd bp[KS], bpi[KS];
    auto gap_entry = [&](int64_t m, int s) {
        const int k = 4 * s + kq, kk = k < n ? k : n - 1;
        const int64_t g = m * nn + kk + n * cc;
        bp[s] = *(col ? e0 + kk : src.P + g);
        bpi[s] = *(col ? e0 + kk : src.Pinv + g);
    };
    // a disk's entry s in two halves: the loads (G, inv(G) into bp, bpi; the row phase, or e0's in column n,
    // into ap), issued with the gap loads' timing, and the products, formed after the region's MFMAs
    zd ap[KS];
    auto disk_load = [&](int64_t m, int s) {
        const int k = 4 * s + kq, kk = k < n ? k : n - 1;
        ap[s] = *(col ? e0 + kk : src.ph + m * n + kk);
        bp[s] = src.G[kk + n * cc];
        bpi[s] = src.Ginv[kk + n * cc];
    };
    auto disk_form = [&](int64_t m) {
        const zd phi = src.phinv[m * n + cc];
#pragma unroll
        for (int s = 0; s < KS; s++) {
            const zd p = cmul(ap[s], bp[s]), pi = cmul(bpi[s], phi);
            bp[s] = cmk<double>(col ? ap[s].re : p.re, col ? ap[s].im : p.im);
            bpi[s] = cmk<double>(col ? ap[s].re : pi.re, col ? ap[s].im : pi.im);
        }
    };
    auto disk_entries = [&](int64_t m) {
#pragma unroll
        for (int s = 0; s < KS; s++) disk_load(m, s);
        disk_form(m);
    };
    if (job.nsteps > 0) {
        const ChainStep s0 = steps[job.step0];
        const int64_t m = job.pbase + s0.reg;
        if (s0.disk) {
            disk_entries(m);
        } else {
#pragma unroll
            for (int s = 0; s < KS; s++) gap_entry(m, s);
        }
    }
    __syncthreads();
    for (int q = 0; q < job.nsteps; q++) {
        const ChainStep st = steps[job.step0 + q];
        const bool more = q + 1 < job.nsteps;
        const ChainStep sn = more ? steps[job.step0 + q + 1] : st;
        const int64_t mn = job.pbase + sn.reg;
        const bool gap_next = more && !sn.disk;
        const zd* U = buf + 2 * (q & 1) * SZ;
        const zd* V = U + SZ;
        zd* Un = buf + 2 * ((q + 1) & 1) * SZ;
        zd* Vn = Un + SZ;
        d4 c21r = {0.0, 0.0, 0.0, 0.0}, c21i = c21r, c22r = c21r, c22i = c21r;
        d4 g21 = c21r, g22 = c21r;  // GAUSS: sums of (Ar + Ai)(Br + Bi); c..r, c..i then hold Ar Br, Ai Bi
#pragma unroll
        for (int s = 0; s < KS; s++) {
            if (KF == 0 && 4 * s >= kp) break;  // k >= n are zero padding
            const int k = 4 * s + kq;
            const zd u = U[ar + LD * k], v = V[ar + LD * k];
            const zd p = bp[s], pi = bpi[s];
            if constexpr (GAUSS) {
                c21r = __builtin_amdgcn_mfma_f64_16x16x4f64(u.re, p.re, c21r, 0, 0, 0);
                c21i = __builtin_amdgcn_mfma_f64_16x16x4f64(u.im, p.im, c21i, 0, 0, 0);
                g21 = __builtin_amdgcn_mfma_f64_16x16x4f64(u.re + u.im, p.re + p.im, g21, 0, 0, 0);
                c22r = __builtin_amdgcn_mfma_f64_16x16x4f64(v.re, pi.re, c22r, 0, 0, 0);
                c22i = __builtin_amdgcn_mfma_f64_16x16x4f64(v.im, pi.im, c22i, 0, 0, 0);
                g22 = __builtin_amdgcn_mfma_f64_16x16x4f64(v.re + v.im, pi.re + pi.im, g22, 0, 0, 0);
            } else {
                c21r = __builtin_amdgcn_mfma_f64_16x16x4f64(u.re, p.re, c21r, 0, 0, 0);
                c21i = __builtin_amdgcn_mfma_f64_16x16x4f64(u.re, p.im, c21i, 0, 0, 0);
                c22r = __builtin_amdgcn_mfma_f64_16x16x4f64(v.re, pi.re, c22r, 0, 0, 0);
                c22i = __builtin_amdgcn_mfma_f64_16x16x4f64(v.re, pi.im, c22i, 0, 0, 0);
                c21r = __builtin_amdgcn_mfma_f64_16x16x4f64(-u.im, p.im, c21r, 0, 0, 0);
                c21i = __builtin_amdgcn_mfma_f64_16x16x4f64(u.im, p.re, c21i, 0, 0, 0);
                c22r = __builtin_amdgcn_mfma_f64_16x16x4f64(-v.im, pi.im, c22r, 0, 0, 0);
                c22i = __builtin_amdgcn_mfma_f64_16x16x4f64(v.im, pi.re, c22i, 0, 0, 0);
            }
            if (gap_next) gap_entry(mn, s);
            if (more && sn.disk) disk_load(mn, s);
        }
        if (more && sn.disk) disk_form(mn);
        if constexpr (GAUSS) {  // re = Ar Br - Ai Bi, im = (Ar + Ai)(Br + Bi) - Ar Br - Ai Bi
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const double a1 = c21r[v], a2 = c21i[v], b1 = c22r[v], b2 = c22i[v];
                c21r[v] = a1 - a2;
                c21i[v] = g21[v] - a1 - a2;
                c22r[v] = b1 - b2;
                c22i[v] = g22[v] - b1 - b2;
            }
        }
        // D layout: register v of lane holds D[4 v + lane / 16][lane % 16] of the tile
        if (own) {
#pragma unroll
            for (int v = 0; v < 4; v++)
                abr[v] = cadd(abr[v], cmul(st.S, cmk<double>(c21r[v] + c22r[v], c21i[v] + c22i[v])));
        }
        if (!spare && wave == 0 && lane < n) {
            zd sacc = cmk<double>(0.0, 0.0);
            for (int k = 0; k < n; k++) sacc = cmac(sacc, cadd(U[lane + LD * k], V[lane + LD * k]), e0[k]);
            abr[0] = cadd(abr[0], cmul(st.S, sacc));
        }
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int i = tr * 16 + 4 * v + (lane >> 4), j = bc;
            if (i < n && j < n) {
                const zd o21 = cmk<double>(c21r[v], c21i[v]), o22 = cmk<double>(c22r[v], c22i[v]);
                zd t21 = o21, t22 = o22;
                // a T21 + b T22 and b T21 + a T22 with one complex product: a + b = 1 at every boundary
                // (a = (nr + nl) / 2nr, b = (nr - nl) / 2nr), the same up to rounding
                if (more) {
                    const zd ad = cmul(sn.a, csub(o21, o22));
                    t21 = cadd(o22, ad);
                    t22 = csub(o21, ad);
                }
                Un[i + LD * j] = t21;
                Vn[i + LD * j] = t22;
            }
        }
        __syncthreads();
    }
    if (own) {
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int i = tr * 16 + 4 * v + (lane >> 4);
            if (i < n) ab[i] = abr[v];
        }
    }
    if (!spare && wave == 0 && lane < n) ab[lane] = abr[0];
    // T21, T22 (the last region's products, in buffer nsteps & 1) compacted to n x n (column-major) over
    // the other buffer for the solves
    const zd* Uf = buf + 2 * (job.nsteps & 1) * SZ;
    zd* T21 = buf + 2 * ((job.nsteps + 1) & 1) * SZ;
    zd* T22 = T21 + SZ;
    for (int e = tid; e < nn; e += NT) {
        const int i = e % n, j = e / n;
        T21[e] = Uf[i + LD * j];
        T22[e] = Uf[SZ + i + LD * j];
    }
    __syncthreads();
    // boost = -T22 \ axion_beam ;  refl = -T22 \ (T21 * reflect)
    if (reflect)
        for (int i = tid; i < n; i += NT) {
            zd s = cmk<double>(0.0, 0.0);
            for (int k = 0; k < n; k++) s = cadd(s, cmul(T21[i + n * k], reflect[k]));
            tmp[n + i] = s;
        }
    __syncthreads();
    lu_factor_block(T22, n, piv, sval, sidx);
    if (tid == 0) {
        for (int i = 0; i < n; i++) tmp[i] = ab[i];
        lu_solve_col(T22, n, piv, tmp);
        for (int i = 0; i < n; i++) boost[(size_t)job.out + i] = cmk<double>(-tmp[i].re, -tmp[i].im);
        if (reflect) {
            lu_solve_col(T22, n, piv, tmp + n);
            for (int i = 0; i < n; i++) refl[(size_t)job.out + i] = cmk<double>(-tmp[n + i].re, -tmp[n + i].im);
        }
    }
}

template <int NP, bool GAUSS, int KF = 0>
static void launch_chain_mfma(hipStream_t s, size_t nj, const ChainJob* jobs, const ChainStep* steps,
                              const ChainSrc& src, int n, const zd* e0, const zd* reflect, zd* boost, zd* refl) {
    if (KF > 0 && (n + 3) / 4 != KF) throw Error(BF_ERR_UNSUPPORTED, "chain kernel k steps do not match n");
    const size_t smem = ChainMfma<NP>::lds_bytes(n);
    allow_lds((const void*)k_chain_mfma<NP, GAUSS, KF>, smem);
    hipLaunchKernelGGL((k_chain_mfma<NP, GAUSS, KF>), dim3((unsigned)nj), dim3(ChainMfma<NP>::NT), smem, s, jobs,
                       steps, src, n, e0, reflect, boost, refl);
}

// The matrix-core chain with the two products in separate waves: wave w of 2 (NP / 16)^2 computes
// output tile w % (NP / 16)^2 of T21' = U P (w in the first half) or T22' = V inv(P) (second half), three
// MFMAs per complex 4-k step.  LDS holds T21 and T22 themselves; each wave forms its A operand
// U = a T21 + b T22 or V = b T21 + a T22 as it reads it, so a region's two barriers are shared by twice
// the waves of k_chain_mfma (four per SIMD with two jobs per CU).  The axion sum S (U + V) e0 comes as
// column n of both products (each half adds its own term); padding, layout and prefetch as k_chain_mfma.
template <int NP>
struct ChainMfma2 {
    static constexpr int TW = NP / 16;
    static constexpr int NTH = 64 * TW * TW;  // threads per product
    static constexpr int NT = 2 * NTH;
    static constexpr int LD = NP + 1;
    static constexpr int PPT = NP * NP / NTH;  // entries of the thread's matrix per region
    static size_t lds_bytes(int n) {
        return sizeof(zd) * ((size_t)4 * NP * LD + 4 * (size_t)n) + (sizeof(double) + sizeof(int)) * NT +
               sizeof(int) * (size_t)n;
    }
};

template <int NP>
__global__ void __launch_bounds__(ChainMfma2<NP>::NT, 4) k_chain_mfma2(const ChainJob* __restrict__ jobs,
                                                                      const ChainStep* __restrict__ steps,
                                                                      const ChainSrc src, int n,
                                                                      const zd* __restrict__ e0_all,
                                                                      const zd* __restrict__ reflect, zd* boost,
                                                                      zd* refl) {
    using T = ChainMfma2<NP>;
    constexpr int LD = T::LD, NT = T::NT, NTH = T::NTH, TW = T::TW;
    constexpr int SZ = NP * LD;
    extern __shared__ unsigned char smem_raw[];
    zd* T21 = reinterpret_cast<zd*>(smem_raw);  // T21[i + LD k]
    zd* T22 = T21 + SZ;
    zd* Pt = T22 + SZ;  // Pt[j + LD k] = P[k][j]
    zd* Pit = Pt + SZ;
    zd* ab = Pit + SZ;  // the two products' axion terms, [2][n]
    zd* tmp = ab + 2 * n;
    double* sval = reinterpret_cast<double*>(tmp + 2 * n);
    int* sidx = reinterpret_cast<int*>(sval + NT);
    int* piv = sidx + NT;
    const ChainJob job = jobs[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int prod = wave / (TW * TW), tile = wave % (TW * TW);  // prod 0: T21' = U P, 1: T22' = V inv(P)
    const int tr = tile % TW, tc = tile / TW;
    const int nn = n * n;
    const bool spare = n < NP;
    for (int e = tid; e < SZ; e += NT) {
        const int i = e % LD, k = e / LD;
        T21[e] = cmk<double>(0.0, 0.0);
        T22[e] = cmk<double>(i == k && i < n ? 1.0 : 0.0, 0.0);
        Pt[e] = cmk<double>(0.0, 0.0);
        Pit[e] = cmk<double>(0.0, 0.0);
    }
    for (int e = tid; e < 2 * n; e += NT) ab[e] = cmk<double>(0.0, 0.0);
    __syncthreads();
    if (spare)
        for (int k = tid; k < n; k += NT) Pt[n + LD * k] = Pit[n + LD * k] = e0_all[job.e0off + k];
    // the thread's entries g = pt + NTH m of its product's matrix (P or inv(P)), stored transposed
    const int pt = tid % NTH;
    zd* Bl = prod ? Pit : Pt;
    auto fetch = [&](const ChainStep& sq, bool ok, zd* dp) {
#pragma unroll
        for (int m = 0; m < T::PPT; m++) {
            const int g = pt + NTH * m;
            dp[m] = cmk<double>(0.0, 0.0);
            if (ok && g < nn) dp[m] = chain_entry(src, sq.disk, job.pbase + sq.reg, n, g, g % n, g / n, prod != 0);
        }
    };
    auto store = [&](const zd* dp) {
#pragma unroll
        for (int m = 0; m < T::PPT; m++) {
            const int g = pt + NTH * m;
            if (g < nn) Bl[g / n + LD * (g % n)] = dp[m];
        }
    };
    if (job.nsteps > 0) {
        zd d0[T::PPT];
        fetch(steps[job.step0], true, d0);
        store(d0);
    }
    zd abr[4];
#pragma unroll
    for (int v = 0; v < 4; v++) abr[v] = cmk<double>(0.0, 0.0);
    const bool own = spare && tc == n / 16 && (lane & 15) == n % 16;  // lanes holding column n
    const int ar = tr * 16 + (lane & 15), bc = tc * 16 + (lane & 15), kq = lane >> 4;
    const int kp = (n + 3) & ~3;  // k steps cover k < n
    __syncthreads();
    for (int q = 0; q < job.nsteps; q++) {
        const ChainStep st = steps[job.step0 + q];
        const bool more = q + 1 < job.nsteps;
        zd pn[T::PPT];  // the next region's entries, loaded while the products run
        fetch(steps[job.step0 + (more ? q + 1 : q)], more, pn);
        // this half's mix of T21, T22: U = a T21 + b T22 (prod 0) or V = b T21 + a T22 (prod 1)
        const zd ma = prod ? st.b : st.a, mb = prod ? st.a : st.b;
        d4 cr = {0.0, 0.0, 0.0, 0.0}, ci = cr, cg = cr;
#pragma unroll
        for (int ks = 0; ks < NP; ks += 4) {
            if (ks >= kp) break;  // k >= n are zero padding
            const int k = ks + kq;
            const zd t1 = T21[ar + LD * k], t2 = T22[ar + LD * k];
            const zd x = cadd(cmul(ma, t1), cmul(mb, t2));
            const zd bo = Bl[bc + LD * k];
            cr = __builtin_amdgcn_mfma_f64_16x16x4f64(x.re, bo.re, cr, 0, 0, 0);
            ci = __builtin_amdgcn_mfma_f64_16x16x4f64(x.im, bo.im, ci, 0, 0, 0);
            cg = __builtin_amdgcn_mfma_f64_16x16x4f64(x.re + x.im, bo.re + bo.im, cg, 0, 0, 0);
        }
        d4 oref, oimf;  // re = Xr Br - Xi Bi, im = (Xr + Xi)(Br + Bi) - Xr Br - Xi Bi
#pragma unroll
        for (int v = 0; v < 4; v++) {
            oref[v] = cr[v] - ci[v];
            oimf[v] = cg[v] - cr[v] - ci[v];
        }
        // D layout: register v of lane holds D[4 v + lane / 16][lane % 16] of the tile
        if (own) {
#pragma unroll
            for (int v = 0; v < 4; v++) abr[v] = cadd(abr[v], cmul(st.S, cmk<double>(oref[v], oimf[v])));
        }
        if (!spare && wave == 0 && lane < n) {
            zd sacc = cmk<double>(0.0, 0.0);
            for (int k = 0; k < n; k++)
                sacc = cmac(sacc, cadd(cadd(cmul(st.a, T21[lane + LD * k]), cmul(st.b, T22[lane + LD * k])),
                                       cadd(cmul(st.b, T21[lane + LD * k]), cmul(st.a, T22[lane + LD * k]))),
                            e0_all[job.e0off + k]);
            abr[0] = cadd(abr[0], cmul(st.S, sacc));
        }
        __syncthreads();
        zd* Tout = prod ? T22 : T21;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int i = tr * 16 + 4 * v + (lane >> 4), j = bc;
            if (i < n && j < n) Tout[i + LD * j] = cmk<double>(oref[v], oimf[v]);
        }
        if (more) store(pn);
        __syncthreads();
    }
    if (own) {
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int i = tr * 16 + 4 * v + (lane >> 4);
            if (i < n) ab[prod * n + i] = abr[v];
        }
    }
    if (!spare && wave == 0 && lane < n) ab[lane] = abr[0];
    __syncthreads();
    // T21, T22 compacted to n x n (column-major) over the P tables for the solves
    zd* C21 = Pt;
    zd* C22 = Pit;
    for (int e = tid; e < nn; e += NT) {
        const int i = e % n, j = e / n;
        C21[e] = T21[i + LD * j];
        C22[e] = T22[i + LD * j];
    }
    __syncthreads();
    // boost = -T22 \ axion_beam ;  refl = -T22 \ (T21 * reflect)
    if (reflect)
        for (int i = tid; i < n; i += NT) {
            zd s = cmk<double>(0.0, 0.0);
            for (int k = 0; k < n; k++) s = cadd(s, cmul(C21[i + n * k], reflect[k]));
            tmp[n + i] = s;
        }
    __syncthreads();
    lu_factor_block(C22, n, piv, sval, sidx);
    if (tid == 0) {
        for (int i = 0; i < n; i++) tmp[i] = cadd(ab[i], ab[n + i]);
        lu_solve_col(C22, n, piv, tmp);
        for (int i = 0; i < n; i++) boost[(size_t)job.out + i] = cmk<double>(-tmp[i].re, -tmp[i].im);
        if (reflect) {
            lu_solve_col(C22, n, piv, tmp + n);
            for (int i = 0; i < n; i++) refl[(size_t)job.out + i] = cmk<double>(-tmp[n + i].re, -tmp[n + i].im);
        }
    }
}

template <int NP>
static void launch_chain_mfma2(hipStream_t s, size_t nj, const ChainJob* jobs, const ChainStep* steps,
                               const ChainSrc& src, int n, const zd* e0, const zd* reflect, zd* boost, zd* refl) {
    const size_t smem = ChainMfma2<NP>::lds_bytes(n);
    allow_lds((const void*)k_chain_mfma2<NP>, smem);
    hipLaunchKernelGGL(k_chain_mfma2<NP>, dim3((unsigned)nj), dim3(ChainMfma2<NP>::NT), smem, s, jobs, steps, src, n,
                       e0, reflect, boost, refl);
}

// the LDS chain for n <= CHAIN_LDS_MAX: one block per job
static void launch_chain_lds(hipStream_t s, size_t nj, const ChainJob* jobs, const ChainStep* steps,
                             const ChainSrc& src, int n, const zd* e0, const zd* reflect, zd* boost, zd* refl) {
    const size_t smem = chain_lds_bytes(n);
    auto go = [&](const void* fn, auto kern) {
        allow_lds(fn, smem);
        hipLaunchKernelGGL(kern, dim3((unsigned)nj), dim3(256), smem, s, jobs, steps, src, n, e0, reflect, boost,
                           refl);
    };
#define BF_CHAIN_CASE(NM, NC)                                         \
    case NC:                                                          \
        go((const void*)k_chain_lds<NM, NC>, k_chain_lds<NM, NC>);       \
        break
    const int nc = chain_cols(n);
    // matrix cores where the padding to 16 or 32 wastes little (measured on the sweep's Bessel sets: n = 15
    // runs fastest with the products in separate waves, n = 25 with both products per wave, n = 9 on the
    // register-tiled VALU kernel); options 4 and 2 / 3 force the MFMA forms for every n <= 32
    const int mf = opts().chain_mfma;
    const bool mfma_pays = (n > 12 && n <= 16) || (n > 24 && n <= 32);
    if (n <= 32 && (mf == 4 || (mf == 1 && n > 12 && n <= 16))) {
        if (n <= 16)
            launch_chain_mfma2<16>(s, nj, jobs, steps, src, n, e0, reflect, boost, refl);
        else
            launch_chain_mfma2<32>(s, nj, jobs, steps, src, n, e0, reflect, boost, refl);
        BF_CHECK_LAUNCH();
        return;
    }
    if (mf == 1 && n > 24 && n <= 32) {  // the sweep's n = 25: k steps fixed at compile time
        if (n <= 28)
            launch_chain_mfma<32, true, 7>(s, nj, jobs, steps, src, n, e0, reflect, boost, refl);
        else
            launch_chain_mfma<32, true, 8>(s, nj, jobs, steps, src, n, e0, reflect, boost, refl);
        BF_CHECK_LAUNCH();
        return;
    }
    if (n <= 32 && (mf == 2 || mf == 3 || (mf == 1 && mfma_pays))) {
        const bool gauss = mf != 2;
        if (n <= 16)
            gauss ? launch_chain_mfma<16, true>(s, nj, jobs, steps, src, n, e0, reflect, boost, refl)
                  : launch_chain_mfma<16, false>(s, nj, jobs, steps, src, n, e0, reflect, boost, refl);
        else
            gauss ? launch_chain_mfma<32, true>(s, nj, jobs, steps, src, n, e0, reflect, boost, refl)
                  : launch_chain_mfma<32, false>(s, nj, jobs, steps, src, n, e0, reflect, boost, refl);
        BF_CHECK_LAUNCH();
        return;
    }
    switch (opts().chain_tile ? n : 0) {  // the Bessel sets of the disk-position sweep (9, 15, 25), 45 LG modes
        case 9:
            launch_chain_tile<9, 1, 2, 64>(s, nj, jobs, steps, src, e0, reflect, boost, refl);
            break;
        case 15:
            launch_chain_tile<15, 2, 2, 64>(s, nj, jobs, steps, src, e0, reflect, boost, refl);
            break;
        case 25:
            launch_chain_tile<25, 2, 3, 128>(s, nj, jobs, steps, src, e0, reflect, boost, refl);
            break;
        case 45:
            launch_chain_tile<45, 3, 3, 256>(s, nj, jobs, steps, src, e0, reflect, boost, refl);
            break;
        default:
            if (n <= 16) {
                go((const void*)k_chain_lds<16, 1>, k_chain_lds<16, 1>);  // floor(256 / n) >= n: one column
            } else if (n <= 32) {
                switch (nc) {
                    BF_CHAIN_CASE(32, 2);
                    BF_CHAIN_CASE(32, 3);
                    BF_CHAIN_CASE(32, 4);
                    default: throw Error(BF_ERR_UNSUPPORTED, "chain column count");
                }
            } else {
                switch (nc) {  // n = 33..48: 5, 6, 7, 9 or 10 columns
                    BF_CHAIN_CASE(CHAIN_LDS_MAX, 5);
                    BF_CHAIN_CASE(CHAIN_LDS_MAX, 6);
                    BF_CHAIN_CASE(CHAIN_LDS_MAX, 7);
                    BF_CHAIN_CASE(CHAIN_LDS_MAX, 9);
                    BF_CHAIN_CASE(CHAIN_LDS_MAX, 10);
                    default: throw Error(BF_ERR_UNSUPPORTED, "chain column count");
                }
            }
    }
#undef BF_CHAIN_CASE
    BF_CHECK_LAUNCH();
}

// disk-region diagonal propagation: P[f][reg] = diag(exp(-1im*kz_m*dz)) * X with
// X = the shared Gram matrix G (flat untilted disk, from_gram[q] != 0) or P[f][reg] itself
// With Pinv and Ginv = inv(G) given, a disk taken from G also gets its inverse,
// inv(diag(ph) G) = inv(G) diag(1 / ph): Pinv[slot] = Ginv with column j divided by ph_j.
__global__ void k_disk_diag(zd* P, const int* __restrict__ slots, const int* __restrict__ from_gram,
                            const zd* __restrict__ G, const zd* __restrict__ k0sq, const double* dz,
                            const zd* __restrict__ kt, int n, zd* Pinv, const zd* __restrict__ Ginv) {
    __shared__ zd sph[BF_MAX_MODES];
    const int q = blockIdx.x;
    const zd k2 = k0sq[q];
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const zd kz = csqrt_j(csub(k2, cmul(kt[i], kt[i])));
        sph[i] = cexp_j(rmul(dz[q], neg_i(kz)));
    }
    __syncthreads();
    const size_t off = (size_t)slots[q] * n * n;
    zd* Pm = P + off;
    const bool fg = from_gram[q] != 0;
    const zd* X = fg ? G : Pm;
    const bool inv = fg && Pinv;
    each2d(threadIdx.x, blockDim.x, n, n, [&](int j, int i) {
        const int e = i + n * j;
        Pm[e] = cmul(sph[i], X[e]);
        if (inv) {
            const zd a = Ginv[e], d = sph[j];
            const double den = d.re * d.re + d.im * d.im;
            Pinv[off + e] = cmk<double>((a.re * d.re + a.im * d.im) / den, (a.im * d.re - a.re * d.im) / den);
        }
    });
}

// P[slot] = s * G  (1D propagator in an air region: prop multiplies by a scalar phase)
__global__ void k_scaled_copy(zd* P, const int* __restrict__ slots, const zd* __restrict__ s, const zd* __restrict__ G,
                              int nn) {
    const int q = blockIdx.x;
    for (int e = threadIdx.x; e < nn; e += blockDim.x) P[(size_t)slots[q] * nn + e] = cmul(G[e], s[q]);
}

// Parseval regions by k-bin group.  H[g] = sum over the group's bins k and the
// components c of conj(A_i(c,k)) B_j(c,k), column-major n x n (i + n j).
__global__ void k_group_gram(const zd* __restrict__ A, const zd* __restrict__ B, size_t K, int C, int n,
                             const int* __restrict__ off, const int* __restrict__ bins, zd* H) {
    const int g = blockIdx.x;
    const int nn = n * n;
    const int b0 = off[g], b1 = off[g + 1];
    const size_t Ktot = (size_t)C * K;
    for (int e = threadIdx.x; e < nn; e += blockDim.x) {
        const int i = e % n, j = e / n;
        double re = 0.0, im = 0.0;
        for (int q = b0; q < b1; q++) {
            const size_t k = (size_t)bins[q];
            for (int c = 0; c < C; c++) {
                const zd a = A[(size_t)i * Ktot + (size_t)c * K + k], b = B[(size_t)j * Ktot + (size_t)c * K + k];
                re += a.re * b.re + a.im * b.im;
                im += a.re * b.im - a.im * b.re;
            }
        }
        H[(size_t)g * nn + e] = cmk<double>(re, im);
    }
}

// The same with the group's pattern values staged in LDS first (each read once
// instead of once per matrix entry); the sums run in the same order.  Dynamic
// LDS: 2 n (bins x C) complex.
__global__ void __launch_bounds__(256) k_group_gram_lds(const zd* __restrict__ A, const zd* __restrict__ B, size_t K,
                                                        int C, int n, const int* __restrict__ off,
                                                        const int* __restrict__ bins, zd* H) {
    extern __shared__ unsigned char smem_raw[];
    const int g = blockIdx.x;
    const int nn = n * n;
    const int b0 = off[g], b1 = off[g + 1], np = (b1 - b0) * C;
    zd* sa = reinterpret_cast<zd*>(smem_raw);  // [i][pair], pair = (q - b0) * C + c
    zd* sb = sa + (size_t)n * np;
    const size_t Ktot = (size_t)C * K;
    each2d(threadIdx.x, blockDim.x, n, np, [&](int i, int p) {
        const size_t k = (size_t)bins[b0 + p / C] + (size_t)(p % C) * K;
        sa[i * np + p] = A[(size_t)i * Ktot + k];
        sb[i * np + p] = B[(size_t)i * Ktot + k];
    });
    __syncthreads();
    each2d(threadIdx.x, blockDim.x, n, n, [&](int j, int i) {
        double re = 0.0, im = 0.0;
        const zd* a_ = sa + i * np;
        const zd* b_ = sb + j * np;
        for (int p = 0; p < np; p++) {
            const zd a = a_[p], b = b_[p];
            re += a.re * b.re + a.im * b.im;
            im += a.re * b.im - a.im * b.re;
        }
        H[(size_t)g * nn + i + n * j] = cmk<double>(re, im);
    });
}

// Flat air regions by k-bin groups on the FP64 matrix cores:
//     P[m] = sum_g W[m][g] H_g,   W[m][g] = exp(-1im*kz*dz_m) * scale,
// kz = sqrt(k0_m^2 - kx_g^2 - ky_g^2) at the group's representative bin (the
// phase of the propagator's per-bin tables, kz_phase).  A block computes a
// BM x BN tile of [problems m] x [matrix entries e = i + n j] and forms its
// rows' weights in registers (never stored in HBM).  Consecutive problems are
// mostly one geometry's air gaps at one frequency, so a block's rows share few
// k0^2 values: the host lists them per block (slots), kz is formed once per
// (slot, group) two chunks ahead into a three-deep LDS table, and each weight
// then costs only exp(-1im*kz*dz) (the same operations as kz_phase: the
// weights are identical).  Per chunk of KC groups the block loads the chunk
// of H (H[g][e], shared by every problem of the mode set), stages weights and
// H in LDS, double-buffered (one barrier per chunk); the MFMAs of a chunk run
// between the next chunk's H loads (issued first) and its weights (formed
// after).  v_mfma_f64_16x16x4f64 with two accumulators per complex 16 x 16
// tile (re += Wre Hre + (-Wim) Him, im += Wre Him + Wim Hre); each wave owns
// 32 problems x 64 entries (2 x 4 tiles) on zero-padded operands (the block
// waits for its fullest wave anyway).  Blocks of one column slab are
// consecutive in the grid, so the blocks in flight read the same H columns.
// Results go straight to each problem's matrix (column-major n x n).
// TAB = true: the weights are read from a table wtab[g][m] (M rows) instead
// (the coefficient rows of the spacing series below); kz and the per-block
// k0^2 lists are then unused.
template <int WN, bool HERM = false, bool TAB = false>
struct PgShape {
    static constexpr int WM = 2;                     // waves along problems
    static constexpr int CT = HERM ? 2 : 4;          // 16-entry column tiles per wave
    static constexpr int BM = 32 * WM, BN = 16 * CT * WN;  // block tile
    // groups per chunk: 16 for the half-width tiles of the Hermitian product on a weight table (as many
    // MFMAs per barrier as the full product's 8; the direct product's kz staging keeps it at 8)
    static constexpr int KC = HERM && TAB ? 16 : 8;
    static constexpr int NT = 64 * WM * WN;
    static constexpr int LDA = BM + 2, LDB = BN + 2;  // padded LDS rows (doubles)
    static constexpr int WPT = BM * KC / NT;          // weights formed per thread and chunk
    static constexpr int HPT = BN * KC / NT;          // H elements loaded per thread and chunk
    static constexpr int KZPT = (BM * KC + NT - 1) / NT;  // (slot, group) pairs per thread at most
    static_assert(NT % BM == 0 && NT % BN == 0 && KC % 4 == 0 && WPT >= 1, "Parseval GEMM thread mapping");
};
constexpr int PG_BM = 64;  // rows per block (all shapes)

// HERM: H holds the packed upper triangles (ncol = n (n + 1) / 2 columns, tri[t] = i | j << 16): with H_g
// Hermitian, A = sum Re(w) H and B = sum Im(w) H are Hermitian and the product is A + i B, so the four
// real sums Wr Hr, Wr Hi, Wi Hr, Wi Hi at (i, j), i <= j, give both (i, j) = (A + i B)_ij and
// (j, i) = conj(A_ij) + i conj(B_ij): half the entries for the same MFMAs per entry.  Else ncol = nn.
template <int WN, bool TAB, bool HERM>
__global__ void __launch_bounds__((PgShape<WN, HERM, TAB>::NT), 4 / WN) k_parseval_gemm(
    const zd* __restrict__ H, int nn, int NG, const double* __restrict__ gkx, const double* __restrict__ gky,
    const PwDesc* __restrict__ pd, const int* __restrict__ row_slot, const int* __restrict__ blk_nk,
    const zd* __restrict__ blk_k0sq, int M, double scale, zd* const* __restrict__ dst, int nmb,
    const zd* __restrict__ wtab, int ncb, int cps, zd* __restrict__ part, int ncol, int n,
    const int* __restrict__ tri) {
    using S = PgShape<WN, HERM, TAB>;
    constexpr int BM = S::BM, BN = S::BN, KC = S::KC, NT = S::NT, CT = S::CT;
    __shared__ double swr[2][KC][S::LDA], swi[2][KC][S::LDA];
    __shared__ double shr[2][KC][S::LDB], shi[2][KC][S::LDB];
    __shared__ zd skz[TAB ? 1 : 3][TAB ? 1 : BM][TAB ? 1 : KC];  // kz staging (direct product only)
    // block -> (row block, column block, group split); split ks sums the chunks [ks cps, (ks + 1) cps)
    const int ks_ = blockIdx.x / (nmb * ncb), rem = blockIdx.x % (nmb * ncb);
    const int mb = rem % nmb, cb = rem / nmb;
    const int m0 = mb * BM, c0 = cb * BN;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w % S::WM, wn = w / S::WM;
    const int nk = TAB ? 0 : blk_nk[mb];
    const zd* k0s = TAB ? nullptr : blk_k0sq + (size_t)mb * BM;
    // weights: thread t forms row m0 + t % BM, chunk columns t / BM + (NT / BM) j
    const int wrow = t % BM, wk0 = t / BM;
    const bool wvalid = m0 + wrow < M;
    double dz = 0.0;
    int slot = 0;
    if (!TAB && wvalid) {
        dz = pd[m0 + wrow].dz;
        slot = row_slot[m0 + wrow];
    }
    // H: thread t loads entries c0 + t % BN, chunk rows t / BN + (NT / BN) j
    const int he = t % BN, hk0 = t / BN;
    const bool hvalid = c0 + he < ncol;
    zd wreg[S::WPT], hreg[S::HPT];
    auto form_kz = [&](int c) {  // kz of chunk c into skz[c % 3]
        if constexpr (!TAB) {
#pragma unroll
            for (int i = 0; i < S::KZPT; i++) {
                const int idx = t + NT * i, s = idx / KC, k = idx % KC, g = c * KC + k;
                if (s < nk && g < NG) skz[c % 3][s][k] = kz_of(k0s[s], gkx[g], gky[g]);
            }
        }
    };
    auto load_h = [&](int c) {
#pragma unroll
        for (int j = 0; j < S::HPT; j++) {
            const int g = c * KC + hk0 + (NT / BN) * j;
            hreg[j] = (hvalid && g < NG) ? H[(size_t)g * ncol + c0 + he] : cmk<double>(0.0, 0.0);
        }
    };
    auto form_w = [&](int c) {  // reads skz[c % 3]
#pragma unroll
        for (int j = 0; j < S::WPT; j++) {
            const int k = wk0 + (NT / BM) * j, g = c * KC + k;
            zd v = cmk<double>(0.0, 0.0);
            if constexpr (TAB) {
                if (wvalid && g < NG) v = wtab[(size_t)g * M + m0 + wrow];
            } else if (wvalid && g < NG) {
                const zd ph = kz_phase_of(skz[c % 3][slot][k], dz);
                v = cmk<double>(ph.re * scale, ph.im * scale);
            }
            wreg[j] = v;
        }
    };
    auto stage = [&](int c) {
        const int b = c & 1;
#pragma unroll
        for (int j = 0; j < S::WPT; j++) {
            const int k = wk0 + (NT / BM) * j;
            swr[b][k][wrow] = wreg[j].re;
            swi[b][k][wrow] = wreg[j].im;
        }
#pragma unroll
        for (int j = 0; j < S::HPT; j++) {
            const int k = hk0 + (NT / BN) * j;
            shr[b][k][he] = hreg[j].re;
            shi[b][k][he] = hreg[j].im;
        }
    };
    constexpr int NA = HERM ? 4 : 2;  // accumulators per tile: (Re, Im) or (Wr Hr, Wr Hi, Wi Hr, Wi Hi)
    d4 acc[2][CT][NA];
#pragma unroll
    for (int rt = 0; rt < 2; rt++)
#pragma unroll
        for (int ct = 0; ct < CT; ct++)
#pragma unroll
            for (int a = 0; a < NA; a++) acc[rt][ct][a] = d4{0.0, 0.0, 0.0, 0.0};
    const int cbeg = ks_ * cps, cend = min((NG + KC - 1) / KC, cbeg + cps);
    form_kz(cbeg);
    form_kz(cbeg + 1);
    load_h(cbeg);
    __syncthreads();
    form_w(cbeg);
    stage(cbeg);
    __syncthreads();
    for (int c = cbeg; c < cend; c++) {
        const int b = c & 1;
        const bool more = c + 1 < cend;
        if (more) load_h(c + 1);
        form_kz(c + 2);  // skz[(c + 2) % 3] was last read in chunk c - 1's weights, before the previous barrier
#pragma unroll
        for (int ks = 0; ks < KC; ks += 4) {
            // A operand 16 x 4: lane -> (row lane % 16, group ks + lane / 16); B operand 4 x 16: (group, entry)
            const int kk = ks + (lane >> 4);
            double ar[2], ai[2], br[CT], bi[CT];
#pragma unroll
            for (int rt = 0; rt < 2; rt++) {
                ar[rt] = swr[b][kk][wm * 32 + rt * 16 + (lane & 15)];
                ai[rt] = swi[b][kk][wm * 32 + rt * 16 + (lane & 15)];
            }
#pragma unroll
            for (int ct = 0; ct < CT; ct++) {
                br[ct] = shr[b][kk][wn * 16 * CT + ct * 16 + (lane & 15)];
                bi[ct] = shi[b][kk][wn * 16 * CT + ct * 16 + (lane & 15)];
            }
#pragma unroll
            for (int rt = 0; rt < 2; rt++) {
                const double nai = -ai[rt];
#pragma unroll
                for (int ct = 0; ct < CT; ct++) {
                    d4* A = acc[rt][ct];
                    if constexpr (HERM) {
                        A[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[rt], br[ct], A[0], 0, 0, 0);
                        A[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[rt], bi[ct], A[1], 0, 0, 0);
                        A[2] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[rt], br[ct], A[2], 0, 0, 0);
                        A[3] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[rt], bi[ct], A[3], 0, 0, 0);
                    } else {
                        A[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[rt], br[ct], A[0], 0, 0, 0);
                        A[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[rt], bi[ct], A[1], 0, 0, 0);
                        A[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai, bi[ct], A[0], 0, 0, 0);
                        A[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[rt], br[ct], A[1], 0, 0, 0);
                    }
                }
            }
        }
        if (more) {
            form_w(c + 1);  // skz[(c + 1) % 3] was formed in chunk c - 1 (or the prologue)
            stage(c + 1);   // stage buffer (c + 1) & 1 was read in chunk c - 1, before the previous barrier
        }
        __syncthreads();
    }
    // D layout (v_mfma_f64_16x16x4f64): register v of lane holds D[4 v + lane / 16][lane % 16]
    const int r0 = m0 + wm * 32, e0 = c0 + wn * 16 * CT;
#pragma unroll
    for (int rt = 0; rt < 2; rt++) {
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int m = r0 + rt * 16 + 4 * v + (lane >> 4);
            if (m >= M) continue;
            zd* out = part ? part + ((size_t)ks_ * M + m) * nn : dst[m];
#pragma unroll
            for (int ct = 0; ct < CT; ct++) {
                const int e = e0 + ct * 16 + (lane & 15);
                if constexpr (HERM) {
                    if (e < ncol) {
                        const int ij = tri[e], i = ij & 0xffff, j = ij >> 16;
                        const double rr = acc[rt][ct][0][v], ri = acc[rt][ct][1][v];
                        const double ir = acc[rt][ct][2][v], ii = acc[rt][ct][3][v];
                        out[i + n * j] = cmk<double>(rr - ii, ri + ir);
                        if (i != j) out[j + n * i] = cmk<double>(rr + ii, ir - ri);
                    }
                } else {
                    if (e < nn) out[e] = cmk<double>(acc[rt][ct][0][v], acc[rt][ct][1][v]);
                }
            }
        }
    }
}

// dst[m][e] = sum over the splits s (in order) of part[s][m][e]
__global__ void __launch_bounds__(256) k_pgemm_reduce(const zd* __restrict__ part, int nsplit, int M, int nn,
                                                      zd* const* __restrict__ dst) {
    const int e = blockIdx.x * 256 + threadIdx.x, m = blockIdx.y;
    if (e >= nn) return;
    zd acc = part[(size_t)m * nn + e];
    for (int s = 1; s < nsplit; s++) acc = cadd(acc, part[((size_t)s * M + m) * nn + e]);
    dst[m][e] = acc;
}

// Tilted/rough disk regions over a band of frequencies.  The weight of region r
// at frequency f is exp(a_f h(x)) with a_f = -1im*k0(f) and the height map
// h = tx*x + ty*y + surface (src/boost_fractor.jl:145-147 as used by
// propagation_matrix, src/boost_fractor_transformer.jl:212-221).  With a_c at
// the band centre, exp(a_f h) = exp(a_c h) * sum_p ((a_f - a_c) h_s)^p / p! * (h/h_s)^p,
// so P_f = sum_p c_{f,p} M_p with the frequency-independent moments
// M_p = Phi^H diag(exp(a_c h) (h/h_s)^p) Phi; the series is cut where its
// remainder bound falls below 1e-17 (the matrices are O(1)).
struct MomDesc {
    zd a;       // a_c = -1im*k0 at the band centre
    double tx, ty;
    int64_t surf;  // element offset of the surface map, -1: flat
    double hinv;   // 1/h_s
    int p, pad;
};

// max over the grid of |tx*x + ty*y + surface| for each region (one block per region)
__global__ void k_height_max(const MomDesc* __restrict__ d, int nx, int ny, const double* __restrict__ X,
                             const double* __restrict__ Y, const zd* __restrict__ surf, double* out) {
    __shared__ double red[256];
    const MomDesc dd = d[blockIdx.x];
    const size_t K = (size_t)nx * ny;
    double m = 0.0;
    for (size_t i = threadIdx.x; i < K; i += blockDim.x) {
        double hr = dd.tx * X[i % nx] + dd.ty * Y[i / nx], hi = 0.0;
        if (dd.surf >= 0) {
            const zd sv = surf[dd.surf + i];
            hr += sv.re;
            hi = sv.im;
        }
        m = fmax(m, sqrt(hr * hr + hi * hi));
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// moment weights exp(a_c h) (h/h_s)^p, one table per descriptor
__global__ void k_build_moments(zd* out, const MomDesc* __restrict__ d, int nx, int ny, const double* __restrict__ X,
                                const double* __restrict__ Y, const zd* __restrict__ surf) {
    const int tab = blockIdx.y;
    const size_t K = (size_t)nx * ny;
    const MomDesc dd = d[tab];
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < K; i += (size_t)gridDim.x * blockDim.x) {
        zd h = cmk<double>(dd.tx * X[i % nx] + dd.ty * Y[i / nx], 0.0);
        if (dd.surf >= 0) h = cadd(h, surf[dd.surf + i]);
        zd v = cexp_j(cmul(dd.a, h));
        const zd u = rmul(dd.hinv, h);
        for (int q = 0; q < dd.p; q++) v = cmul(v, u);
        out[(size_t)tab * K + i] = v;
    }
}

struct CombJob {
    int slot, q0, np, c0;  // output slot, first moment, moment count, first coefficient
};

// P[slot] = sum_p coef[c0 + p] * M[q0 + p]
__global__ void k_combine_moments(zd* P, const zd* __restrict__ M, const CombJob* __restrict__ jobs,
                                  const zd* __restrict__ coef, int nn) {
    const CombJob jb = jobs[blockIdx.y];
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nn; e += gridDim.x * blockDim.x) {
        double re = 0.0, im = 0.0;
        for (int p = 0; p < jb.np; p++) {
            const zd c = coef[jb.c0 + p], m = M[(size_t)(jb.q0 + p) * nn + e];
            re += c.re * m.re - c.im * m.im;
            im += c.re * m.im + c.im * m.re;
        }
        P[(size_t)jb.slot * nn + e] = cmk<double>(re, im);
    }
}

__global__ void k_copy_z(zd* dst, const zd* __restrict__ src, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// *flag = 1 if a and b differ in any bit (n complex entries)
__global__ void k_differs_z(const zd* __restrict__ a, const zd* __restrict__ b, size_t n, int* flag) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        if (__double_as_longlong(a[i].re) != __double_as_longlong(b[i].re) ||
            __double_as_longlong(a[i].im) != __double_as_longlong(b[i].im))
            *flag = 1;
}

// the upper triangles (i <= j) of the group Grams, packed: Hp[g][t] = H[g][i + n j] with tri[t] = i | j << 16
__global__ void k_pack_upper(const zd* __restrict__ H, int G, int n, const int* __restrict__ tri, int nt,
                             zd* __restrict__ Hp) {
    const size_t total = (size_t)G * nt;
    for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < total; x += (size_t)gridDim.x * blockDim.x) {
        const size_t g = x / nt;
        const int ij = tri[x % nt];
        Hp[x] = H[g * n * n + (ij & 0xffff) + (size_t)n * (ij >> 16)];
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static inline unsigned g1(size_t n, unsigned cap = 2048) {
    size_t b = (n + 255) / 256;
    return (unsigned)std::max<size_t>(1, std::min<size_t>(b, cap));
}
static zd k0_of(zd eps, double lambda) { return rmul((2.0 * M_PI) / lambda, csqrt_j(eps)); }
static zd cdiv_h(zd a, zd b) {
    if (fabs(b.re) >= fabs(b.im)) {
        double r = b.im / b.re, d = b.re + b.im * r;
        return cmk<double>((a.re + a.im * r) / d, (a.im - a.re * r) / d);
    }
    double r = b.re / b.im, d = b.re * r + b.im;
    return cmk<double>((a.re * r + a.im) / d, (a.im * r - a.re) / d);
}


// inv(P[sel[q]]) -> Pinv[sel[q]] for q < count (dsel on the device)
static void launch_inverse(bf_ctx* ctx, const zd* P, zd* Pinv, int n, const int* dsel, size_t count) {
    hipStream_t s = ctx->stream;
    if (count == 0) return;
    if (n <= INV_WAVE_MAX) {
        const size_t smem = sizeof(zd) * (size_t)n * n;
        allow_lds((const void*)k_batched_inverse_wave, smem);
        hipLaunchKernelGGL(k_batched_inverse_wave, dim3((unsigned)count), dim3(64), smem, s, P, Pinv, n, dsel,
                           (const InvSel*)nullptr);
        BF_CHECK_LAUNCH();
    } else {
        DevBuf work(sizeof(zd) * (size_t)n * n * count);
        hipLaunchKernelGGL(k_batched_inverse, dim3((unsigned)count), dim3(256), 0, s, P, Pinv, n, work.as<zd>(), dsel);
        BF_CHECK_LAUNCH();
        BF_HIP(hipStreamSynchronize(s));  // `work` is released on return
    }
}

// Groups of k bins sharing a Parseval weight and their Gram matrices H_g.
// A bin's weight depends on (kx^2, ky^2) of the fftshifted entry it multiplies.
// Squares that agree to 1e-14 relative form one class (grids built from
// floating-point ranges are symmetric only to a few ulp), and the pair of
// classes is unordered.  The group's weight is evaluated at its lowest bin; for
// the others the phase argument then differs at the rounding level (at most
// ~dz*sqrt(|d(k^2)|) ~ 1e-7 rad for a bin sitting exactly on the evanescent
// threshold, far less elsewhere).  Returns false (no grouping) for non-finite k grids.
static bool build_parseval_groups(bf_ctx* ctx, const Grid& g, const ModesDev& md, ModeCache& mc) {
    hipStream_t s = ctx->stream;
    const int nx = g.nx, ny = g.ny, n = md.n;
    const size_t K = g.K(), nn = (size_t)n * n;
    if (!mc.haveIdx) {
        for (double v : g.hkX) if (!std::isfinite(v)) return false;
        for (double v : g.hkY) if (!std::isfinite(v)) return false;
        // classes of squared wavenumbers
        std::vector<double> sq;
        for (double v : g.hkX) sq.push_back(v * v);
        for (double v : g.hkY) sq.push_back(v * v);
        std::sort(sq.begin(), sq.end());
        std::vector<double> cls;  // first (smallest) value of each class
        for (double v : sq)
            if (cls.empty() || v > cls.back() * (1.0 + 1e-14)) cls.push_back(v);
        auto class_of = [&](double v) {
            return (int)(std::upper_bound(cls.begin(), cls.end(), v) - cls.begin()) - 1;
        };
        std::vector<int> cx(nx), cy(ny);
        for (int j = 0; j < nx; j++) cx[j] = class_of(g.hkX[j] * g.hkX[j]);
        for (int j = 0; j < ny; j++) cy[j] = class_of(g.hkY[j] * g.hkY[j]);
        std::vector<std::pair<std::pair<int, int>, int>> keyed(K);
        for (int my = 0; my < ny; my++)
            for (int mx = 0; mx < nx; mx++) {
                int a = cx[(mx + nx / 2) % nx], b = cy[(my + ny / 2) % ny];
                if (b < a) std::swap(a, b);
                keyed[(size_t)my * nx + mx] = {{a, b}, my * nx + mx};
            }
        std::sort(keyed.begin(), keyed.end());
        std::vector<int> off, bins(K);
        std::vector<double> gx, gy;
        double k2max = 0.0;
        for (size_t q = 0; q < K; q++) {
            if (q == 0 || keyed[q].first != keyed[q - 1].first) {
                off.push_back((int)q);
                const int bin = keyed[q].second;  // lowest bin index of the group
                gx.push_back(g.hkX[((bin % nx) + nx / 2) % nx]);
                gy.push_back(g.hkY[((bin / nx) + ny / 2) % ny]);
                k2max = std::max(k2max, gx.back() * gx.back() + gy.back() * gy.back());
            }
            bins[q] = keyed[q].second;
        }
        off.push_back((int)K);
        const int G = (int)gx.size();
        mc.goff.alloc(sizeof(int) * off.size());
        mc.gbin.alloc(sizeof(int) * K);
        upload(mc.goff.p, off.data(), sizeof(int) * off.size(), s);
        upload(mc.gbin.p, bins.data(), sizeof(int) * K, s);
        mc.gkx.alloc(sizeof(double) * G);
        mc.gky.alloc(sizeof(double) * G);
        upload(mc.gkx.p, gx.data(), sizeof(double) * G, s);
        upload(mc.gky.p, gy.data(), sizeof(double) * G, s);
        BF_HIP(hipStreamSynchronize(s));  // the host vectors go out of scope
        mc.ngroups = G;
        mc.gk2max = k2max;
        mc.gmax = 0;
        for (int q = 0; q < G; q++) mc.gmax = std::max(mc.gmax, off[q + 1] - off[q]);
        mc.haveIdx = true;
    }
    const int G = mc.ngroups;
    if (mc.H.bytes < sizeof(zd) * nn * G) mc.H.alloc(sizeof(zd) * nn * G);
    const size_t smem = 2 * sizeof(zd) * (size_t)n * mc.gmax * md.C;
    if (smem <= 64 * 1024) {
        allow_lds((const void*)k_group_gram_lds, smem);
        hipLaunchKernelGGL(k_group_gram_lds, dim3(G), dim3(256), smem, s, mc.Fhat.as<zd>(), mc.FhatM.as<zd>(), K,
                           md.C, n, mc.goff.as<int>(), mc.gbin.as<int>(), mc.H.as<zd>());
    } else {
        hipLaunchKernelGGL(k_group_gram, dim3(G), dim3(256), 0, s, mc.Fhat.as<zd>(), mc.FhatM.as<zd>(), K, md.C, n,
                           mc.goff.as<int>(), mc.gbin.as<int>(), mc.H.as<zd>());
    }
    BF_CHECK_LAUNCH();
    mc.haveH = true;
    // With the masked and unmasked pattern transforms equal (patterns that vanish outside the aperture),
    // H_g = sum conj(Fhat_i) Fhat_j is Hermitian, and so are sum_g Re(w_g) H_g and sum_g Im(w_g) H_g:
    // the product needs the upper triangles only (k_parseval_gemm<.., HERM>)
    mc.herm = false;
    {
        const size_t cnt = (size_t)n * md.Ktot;
        DevBuf& dflag = scratch_buf(ctx, SB_HFLAG, sizeof(int));
        BF_HIP(hipMemsetAsync(dflag.p, 0, sizeof(int), s));
        hipLaunchKernelGGL(k_differs_z, dim3(g1(cnt, 4096)), dim3(256), 0, s, mc.Fhat.as<zd>(), mc.FhatM.as<zd>(),
                           cnt, dflag.as<int>());
        BF_CHECK_LAUNCH();
        int differs = 1;
        download(&differs, dflag.p, sizeof(int), s);
        BF_HIP(hipStreamSynchronize(s));
        if (!differs) {
            std::vector<int> tri;
            for (int j = 0; j < n; j++)
                for (int i = 0; i <= j; i++) tri.push_back(i | j << 16);
            mc.nt = (int)tri.size();
            mc.tri.alloc(sizeof(int) * tri.size());
            upload(mc.tri.p, tri.data(), sizeof(int) * tri.size(), s);
            mc.Hp.alloc(sizeof(zd) * (size_t)G * mc.nt);
            hipLaunchKernelGGL(k_pack_upper, dim3(g1((size_t)G * mc.nt, 4096)), dim3(256), 0, s, mc.H.as<zd>(), G, n,
                               mc.tri.as<int>(), mc.nt, mc.Hp.as<zd>());
            BF_CHECK_LAUNCH();
            BF_HIP(hipStreamSynchronize(s));  // the host table goes out of scope
            mc.herm = true;
        }
    }
    return true;
}

// Weighted Gram launcher: problems share n, Ktot and the weight period Kw.
struct Gram {
    bf_ctx* ctx;
    int n;
    size_t Ktot, Kw;
    DevBuf dprobs, dgroups, partial;
    template <int T>
    void launch_grouped(size_t smem, dim3 grid, int n_, size_t kper) {
        allow_lds((const void*)k_wgram_grouped<T>, smem);
        hipLaunchKernelGGL(k_wgram_grouped<T>, grid, dim3(64 * WgShape<T>::WAVES), smem, ctx->stream,
                           dgroups.as<GemmGroup>(), dprobs.as<GemmProb>(), n_, Ktot, Kw, kper, partial.as<zd>());
        BF_CHECK_LAUNCH();
    }
    void run(const std::vector<GemmProb>& probs, zd* P) {
        if (probs.empty()) return;
        hipStream_t s = ctx->stream;
        const int np = (int)probs.size();
        const int T = (n + 15) / 16;
        const int waves = std::max(1, std::min(16, (T * T + TILES_PER_WAVE - 1) / TILES_PER_WAVE));
        // split the reduction so that the launch has >= ~1024 blocks
        size_t splits = std::max<size_t>(1, (1024 + np - 1) / np);
        splits = std::min(splits, (Ktot + KC - 1) / KC);
        size_t kper = (Ktot + splits - 1) / splits;
        kper = (kper + KC - 1) / KC * KC;
        splits = (Ktot + kper - 1) / kper;
        const size_t nn = (size_t)n * n;
        // grouped kernel for n <= 48 (at T = 4 its double buffer would not fit the registers)
        const int gmax = T == 1 ? WgShape<1>::G : T == 2 ? WgShape<2>::G : T == 3 ? WgShape<3>::G : 1;
        if (T <= 3 && gmax >= 2) {
            // problems sorted so that groups share A and B; a group holds <= gmax of them
            std::vector<int> order(np);
            for (int q = 0; q < np; q++) order[q] = q;
            std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
                return std::make_pair(probs[x].A, probs[x].B) < std::make_pair(probs[y].A, probs[y].B);
            });
            std::vector<GemmProb> sorted(np);
            for (int q = 0; q < np; q++) sorted[q] = probs[order[q]];
            std::vector<GemmGroup> groups;
            for (int q = 0; q < np;) {
                int c = 1;
                while (q + c < np && c < gmax && sorted[q + c].A == sorted[q].A && sorted[q + c].B == sorted[q].B) c++;
                groups.push_back(GemmGroup{sorted[q].A, sorted[q].B, q, c});
                q += c;
            }
            const int ng = (int)groups.size();
            // enough blocks for several rounds of the GPU, so the last round's idle share is small
            size_t gsplits = std::max<size_t>(1, (4096 + ng - 1) / ng);
            gsplits = std::min(gsplits, (Ktot + KC - 1) / KC);
            size_t gkper = (Ktot + gsplits - 1) / gsplits;
            gkper = (gkper + KC - 1) / KC * KC;
            gsplits = (Ktot + gkper - 1) / gkper;
            dprobs.alloc(sizeof(GemmProb) * np);
            upload(dprobs.p, sorted.data(), sizeof(GemmProb) * np, s);
            dgroups.alloc(sizeof(GemmGroup) * ng);
            upload(dgroups.p, groups.data(), sizeof(GemmGroup) * ng, s);
            if (partial.bytes < sizeof(zd) * nn * gsplits * np) partial.alloc(sizeof(zd) * nn * gsplits * np);
            const size_t smem = 2 * sizeof(zd) * ((size_t)T * 16 * LDA * 2 + (size_t)WG_G * KC);
            {
                TimedLaunch tl(ctx, "wgram", 8.0 * (double)nn * (double)Ktot * np);
                switch (T) {
                    case 1: launch_grouped<1>(smem, dim3((unsigned)gsplits, ng), n, gkper); break;
                    case 2: launch_grouped<2>(smem, dim3((unsigned)gsplits, ng), n, gkper); break;
                    default: launch_grouped<3>(smem, dim3((unsigned)gsplits, ng), n, gkper); break;
                }
            }
            hipLaunchKernelGGL(k_wgram_reduce, dim3(g1(nn, 64), np), dim3(256), 0, s, dprobs.as<GemmProb>(),
                               partial.as<zd>(), (int)gsplits, (int)nn, P);
            BF_CHECK_LAUNCH();
            BF_HIP(hipStreamSynchronize(s));  // dprobs/dgroups/partial are reused by the next call
            return;
        }
        dprobs.alloc(sizeof(GemmProb) * np);
        upload(dprobs.p, probs.data(), sizeof(GemmProb) * np, s);
        if (partial.bytes < sizeof(zd) * nn * splits * np) partial.alloc(sizeof(zd) * nn * splits * np);
        const size_t smem = sizeof(zd) * (size_t)T * 16 * LDA * 2;
        allow_lds((const void*)k_wgram, smem);
        {
            // algorithmic work: one complex multiply-add (8 flop) per (i, j, k) and problem
            TimedLaunch tl(ctx, "wgram", 8.0 * (double)nn * (double)Ktot * np);
            const int zb = (T * T + waves * TILES_PER_WAVE - 1) / (waves * TILES_PER_WAVE);
            hipLaunchKernelGGL(k_wgram, dim3((unsigned)splits, np, zb), dim3(64 * waves), smem, s,
                               dprobs.as<GemmProb>(), n, Ktot, Kw, kper, partial.as<zd>());
            BF_CHECK_LAUNCH();
        }
        hipLaunchKernelGGL(k_wgram_reduce, dim3(g1(nn, 64), np), dim3(256), 0, s, dprobs.as<GemmProb>(),
                           partial.as<zd>(), (int)splits, (int)nn, P);
        BF_CHECK_LAUNCH();
        BF_HIP(hipStreamSynchronize(s));  // dprobs/partial are reused by the next call
    }
};

// All propagation matrices P[f][reg] for the regions in `regs` (0-based), at frequencies fr[0..nf).
// eps: the permittivities to use (the transformer passes NaN-fixed values).

// The mode set's frequency-independent tables, formed on first use and kept in mc: the Gram matrix G
// (flat disks, 1D air), the pattern transforms Fhat (masked and unmasked) and, for Parseval regions, the
// k-bin groups and their Grams H.  Returns whether the grouped Parseval product is available.
bool mode_tables(bf_ctx* ctx, const Grid& g, const ModesDev& md, ModeCache& mc, bool need_gram, bool need_parseval,
                 double diskR) {
    hipStream_t s = ctx->stream;
    const int n = md.n;
    const size_t K = g.K(), Ktot = md.Ktot, nn = (size_t)n * n;
    if (need_gram && !mc.haveG) {
        Gram gram{ctx, n, Ktot, K};
        if (mc.G.bytes < sizeof(zd) * nn) mc.G.alloc(sizeof(zd) * nn);
        GemmProb p{md.pat.as<zd>(), md.pat.as<zd>(), nullptr, 0, 0};
        gram.run({p}, mc.G.as<zd>());
        mc.haveG = true;
    }
    if (need_parseval && !mc.haveF) {
        if (mc.Fhat.bytes < sizeof(zd) * n * Ktot) mc.Fhat.alloc(sizeof(zd) * n * Ktot);
        if (mc.FhatM.bytes < sizeof(zd) * n * Ktot) mc.FhatM.alloc(sizeof(zd) * n * Ktot);
        hipLaunchKernelGGL(k_copy_z, dim3(g1(n * Ktot, 4096)), dim3(256), 0, s, mc.Fhat.as<zd>(), md.pat.as<zd>(),
                           n * Ktot);
        hipLaunchKernelGGL(k_copy_z, dim3(g1(n * Ktot, 4096)), dim3(256), 0, s, mc.FhatM.as<zd>(), md.pat.as<zd>(),
                           n * Ktot);
        BF_CHECK_LAUNCH();
        fft_fields(ctx, g, mc.Fhat.as<zd>(), n * md.C, false, diskR);
        fft_fields(ctx, g, mc.FhatM.as<zd>(), n * md.C, true, diskR);
        mc.haveF = true;
    }
    if (!need_parseval || !opts().parseval_groups) return false;
    return mc.haveH || build_parseval_groups(ctx, g, md, mc);
}

void propagation_matrices_core(bf_ctx* ctx, const Grid& g, const ModesDev& md, int kind, double diskR, int nreg,
                               const double* distance, const zd* eps, const double* tx, const double* ty,
                               const zd* surfaces_dev, const std::vector<bool>& flat,
                               const double* fr, int nf, const std::vector<int>& regs, zd* P, ModeCache* cache,
                               ParsevalQueue* queue) {
    hipStream_t s = ctx->stream;
    const int n = md.n;
    const size_t K = g.K(), Ktot = md.Ktot, nn = (size_t)n * n;
    Gram gram{ctx, n, Ktot, K};
    // classify
    std::vector<int> parseval, explicit_air, onedim, disks;
    for (int r : regs) {
        const bool air = (eps[r].re == 1.0);
        if (!air) disks.push_back(r);
        else if (kind == BF_PROP_1D) onedim.push_back(r);
        else if (kind == BF_PROP_NOTILTS || (kind == BF_PROP_PROPAGATOR && flat[r] && tx[r] == 0.0 && ty[r] == 0.0))
            parseval.push_back(r);
        else explicit_air.push_back(r);
    }
    bool need_gram = false;
    for (int r : disks) if (flat[r] && tx[r] == 0.0 && ty[r] == 0.0) need_gram = true;
    if (!onedim.empty()) need_gram = true;
    ModeCache local;
    ModeCache& mc = cache ? *cache : local;
    DevBuf& Fhat = mc.Fhat;
    DevBuf& FhatM = mc.FhatM;
    // Parseval regions through k-bin groups (option parseval_groups=0: per-bin weights)
    const bool grouped = mode_tables(ctx, g, md, mc, need_gram, !parseval.empty(), diskR);
    // frequency-dependent pieces, in frequency batches bounded by table memory
    // tilted/rough disk regions: moments of the frequency expansion (option disk_expansion=0: per-frequency tables)
    std::vector<int> exp_q0(nreg, -1), exp_np(nreg, 0);
    std::vector<double> exp_hs(nreg, 0.0);
    std::vector<zd> exp_ac(nreg);
    DevBuf Mom;
    size_t n_direct_disks = 0;
    {
        std::vector<int> rough;
        for (int r : disks)
            if (!(flat[r] && tx[r] == 0.0 && ty[r] == 0.0)) rough.push_back(r);
        if (!rough.empty() && nf > 2 && opts().disk_expansion) {
            std::vector<MomDesc> hd;
            for (int r : rough)
                hd.push_back(MomDesc{cmk<double>(0.0, 0.0), tx[r], ty[r], flat[r] ? -1 : (int64_t)(K * r), 0.0, 0, 0});
            DevBuf dhd(sizeof(MomDesc) * hd.size()), dhm(sizeof(double) * hd.size());
            upload(dhd.p, hd.data(), sizeof(MomDesc) * hd.size(), s);
            hipLaunchKernelGGL(k_height_max, dim3((unsigned)hd.size()), dim3(256), 0, s, dhd.as<MomDesc>(), g.nx, g.ny,
                               g.X.as<double>(), g.Y.as<double>(), surfaces_dev, dhm.as<double>());
            BF_CHECK_LAUNCH();
            std::vector<double> hmax(hd.size());
            download(hmax.data(), dhm.p, sizeof(double) * hd.size(), s);
            BF_HIP(hipStreamSynchronize(s));
            const double fmin = *std::min_element(fr, fr + nf), fmax = *std::max_element(fr, fr + nf);
            const double fc = 0.5 * (fmin + fmax);
            std::vector<MomDesc> md_;
            for (size_t t = 0; t < rough.size(); t++) {
                const int r = rough[t];
                const double hs = hmax[t];
                if (!std::isfinite(hs)) continue;
                const zd kc = k0_of(eps[r], 299792458.0 / fc);
                const zd ac = neg_i(kc);
                const zd d1 = csub(k0_of(eps[r], 299792458.0 / fmin), kc), d2 = csub(k0_of(eps[r], 299792458.0 / fmax), kc);
                const double delta = std::max(std::hypot(d1.re, d1.im), std::hypot(d2.re, d2.im)) * hs;
                const double growth = std::exp(std::hypot(ac.re, ac.im) * hs + delta);
                int P = 0;
                double tail = growth * delta;  // delta^(P+1)/(P+1)! * e^delta * |exp(a_c h)| bound
                while (tail >= 1e-17 && P < 24) {
                    P++;
                    tail *= delta / (P + 1);
                }
                if (tail >= 1e-17 || P + 1 >= nf) continue;  // the series does not pay (or does not converge fast)
                exp_q0[r] = (int)md_.size();
                exp_np[r] = P + 1;
                exp_hs[r] = hs;
                exp_ac[r] = ac;
                for (int p = 0; p <= P; p++)
                    md_.push_back(MomDesc{ac, tx[r], ty[r], flat[r] ? -1 : (int64_t)(K * r), hs > 0.0 ? 1.0 / hs : 0.0,
                                          p, 0});
            }
            if (!md_.empty()) {
                Mom.alloc(sizeof(zd) * nn * md_.size());
                const size_t chunk = std::max<size_t>(1, (512u << 20) / (K * sizeof(zd)));
                DevBuf tabs(sizeof(zd) * K * std::min(chunk, md_.size())), dmd(sizeof(MomDesc) * md_.size());
                upload(dmd.p, md_.data(), sizeof(MomDesc) * md_.size(), s);
                for (size_t q0 = 0; q0 < md_.size(); q0 += chunk) {
                    const size_t nq = std::min(chunk, md_.size() - q0);
                    hipLaunchKernelGGL(k_build_moments, dim3(g1(K, 1024), (unsigned)nq), dim3(256), 0, s, tabs.as<zd>(),
                                       dmd.as<MomDesc>() + q0, g.nx, g.ny, g.X.as<double>(), g.Y.as<double>(),
                                       surfaces_dev);
                    BF_CHECK_LAUNCH();
                    std::vector<GemmProb> mp;
                    for (size_t q = 0; q < nq; q++)
                        mp.push_back(GemmProb{md.pat.as<zd>(), md.pat.as<zd>(), tabs.as<zd>() + q * K, (int)(q0 + q), 0});
                    gram.run(mp, Mom.as<zd>());
                }
            }
        }
        for (int r : rough)
            if (exp_np[r] == 0) n_direct_disks++;
    }
    const size_t per_f = std::max<size_t>(1, (grouped ? 0 : parseval.size()) + n_direct_disks) * K * sizeof(zd);
    const int fb = (int)std::max<size_t>(
        1, std::min<size_t>({(size_t)nf, (512u << 20) / per_f, 65535 / std::max<size_t>(1, disks.size())}));
    DevBuf ktabs, rtabs, scratch, prop, scr2;
    DevBuf& dslots = ctx->scratch[SB_SLOTS];
    DevBuf& dflag = ctx->scratch[SB_FLAG];
    DevBuf& dk0 = ctx->scratch[SB_K0];
    DevBuf& ddz = ctx->scratch[SB_DZ];
    ParsevalQueue local_q;
    ParsevalQueue& pq = queue ? *queue : local_q;
    // k0 of a permittivity at the batch's frequencies, formed once per distinct permittivity (bitwise)
    std::vector<std::pair<zd, std::vector<zd>>> k0_memo;
    auto k0_at = [&](const zd& e, int f0, int nb) -> const std::vector<zd>& {
        for (auto& m : k0_memo)
            if (std::memcmp(&m.first, &e, sizeof(zd)) == 0 && (int)m.second.size() == nb) return m.second;
        std::vector<zd> v((size_t)nb);
        for (int f = 0; f < nb; f++) v[(size_t)f] = k0_of(e, 299792458.0 / fr[f0 + f]);
        k0_memo.emplace_back(e, std::move(v));
        return k0_memo.back().second;
    };
    for (int f0 = 0; f0 < nf; f0 += fb) {
        const int nb = std::min(fb, nf - f0);
        k0_memo.clear();
        // Parseval air regions: w = exp(-i kz dz)/K in unshifted bins
        std::vector<GemmProb> probs;
        if (grouped) {
            // queued: P[slot] = sum_g W[m][g] H_g, one GEMM for all of them
            pq.pd.reserve(pq.pd.size() + (size_t)nb * parseval.size());
            pq.dst.reserve(pq.dst.size() + (size_t)nb * parseval.size());
            for (int f = 0; f < nb; f++)
                for (int r : parseval) {
                    zd e = eps[r];
                    if (kind == BF_PROP_PROPAGATOR && e.im == 0.0) e.im = -0.0;
                    const zd k0 = k0_at(e, f0, nb)[(size_t)f];
                    pq.pd.push_back(PwDesc{cmul(k0, k0), distance[r], 0, 0});
                    pq.dst.push_back(P + ((size_t)f * nreg + r) * nn);
                }
        } else if (!parseval.empty()) {
            std::vector<KTabDesc> kd;
            for (int f = 0; f < nb; f++)
                for (int r : parseval) {
                    zd e = eps[r];
                    if (kind == BF_PROP_PROPAGATOR && e.im == 0.0) e.im = -0.0;
                    const zd k0 = k0_of(e, 299792458.0 / fr[f0 + f]);
                    kd.push_back(KTabDesc{cmul(k0, k0), distance[r], 0, 0});
                }
            if (ktabs.bytes < sizeof(zd) * K * kd.size()) ktabs.alloc(sizeof(zd) * K * kd.size());
            build_ktabs(ctx, g, kd, 1.0 / (double)((int64_t)g.nx * g.ny), ktabs.as<zd>());
            int q = 0;
            for (int f = 0; f < nb; f++)
                for (int r : parseval) {
                    probs.push_back(GemmProb{Fhat.as<zd>(), FhatM.as<zd>(), ktabs.as<zd>() + (size_t)q * K,
                                             (int)((size_t)f * nreg + r), 0});
                    q++;
                }
        }
        // disk regions: weights = tilt*surface phase, or the shared Gram matrix
        std::vector<int> slots, from_gram;
        std::vector<zd> k0sq;
        std::vector<double> dzs;
        std::vector<CombJob> comb;
        std::vector<zd> coefs;
        {
            std::vector<RTabDesc> rd;
            std::vector<std::pair<int, int>> rd_slot;
            for (int f = 0; f < nb; f++)
                for (int r : disks) {
                    const zd k0 = k0_at(eps[r], f0, nb)[(size_t)f];
                    const int slot = (int)((size_t)f * nreg + r);
                    slots.push_back(slot);
                    k0sq.push_back(cmul(k0, k0));
                    dzs.push_back(distance[r]);
                    from_gram.push_back(flat[r] && tx[r] == 0.0 && ty[r] == 0.0);
                    if (!from_gram.back() && exp_np[r] > 0) {
                        // c_p = ((a_f - a_c) h_s)^p / p!
                        const zd dah = rmul(exp_hs[r], csub(neg_i(k0), exp_ac[r]));
                        comb.push_back(CombJob{slot, exp_q0[r], exp_np[r], (int)coefs.size()});
                        zd c = cmk<double>(1.0, 0.0);
                        for (int p = 0; p < exp_np[r]; p++) {
                            coefs.push_back(c);
                            c = rmul(1.0 / (p + 1), cmul(c, dah));
                        }
                    } else if (!from_gram.back()) {
                        rd.push_back(RTabDesc{neg_i(k0), tx[r], ty[r], flat[r] ? -1 : (int64_t)(K * r)});
                        rd_slot.push_back({slot, (int)rd.size() - 1});
                    }
                }
            if (!rd.empty()) {
                if (rtabs.bytes < sizeof(zd) * K * rd.size()) rtabs.alloc(sizeof(zd) * K * rd.size());
                build_rtabs(ctx, g, rd, surfaces_dev, rtabs.as<zd>());
                for (auto& ps : rd_slot)
                    probs.push_back(GemmProb{md.pat.as<zd>(), md.pat.as<zd>(), rtabs.as<zd>() + (size_t)ps.second * K,
                                             ps.first, 0});
            }
        }
        gram.run(probs, P);
        if (!comb.empty()) {
            DevBuf dcomb(sizeof(CombJob) * comb.size()), dcoef(sizeof(zd) * coefs.size());
            upload(dcomb.p, comb.data(), sizeof(CombJob) * comb.size(), s);
            upload(dcoef.p, coefs.data(), sizeof(zd) * coefs.size(), s);
            hipLaunchKernelGGL(k_combine_moments, dim3(g1(nn, 64), (unsigned)comb.size()), dim3(256), 0, s, P,
                               Mom.as<zd>(), dcomb.as<CombJob>(), dcoef.as<zd>(), (int)nn);
            BF_CHECK_LAUNCH();
            BF_HIP(hipStreamSynchronize(s));
        }
        // diag(exp(-i kz dz)) from the left for the disk regions
        if (!slots.empty() && queue && queue->diag_base) {
            const int64_t b = (int64_t)((P - queue->diag_base) / (ptrdiff_t)nn);
            for (size_t q = 0; q < slots.size(); q++) {
                queue->dslot.push_back((int)(b + slots[q]));
                queue->dfrom.push_back(from_gram[q]);
                queue->dk0sq.push_back(k0sq[q]);
                queue->ddz.push_back(dzs[q]);
            }
        } else if (!slots.empty()) {
            if (dslots.bytes < sizeof(int) * slots.size()) dslots.alloc(sizeof(int) * slots.size());
            if (dflag.bytes < sizeof(int) * slots.size()) dflag.alloc(sizeof(int) * slots.size());
            if (dk0.bytes < sizeof(zd) * slots.size()) dk0.alloc(sizeof(zd) * slots.size());
            if (ddz.bytes < sizeof(double) * slots.size()) ddz.alloc(sizeof(double) * slots.size());
            upload(dslots.p, slots.data(), sizeof(int) * slots.size(), s);
            upload(dflag.p, from_gram.data(), sizeof(int) * slots.size(), s);
            upload(dk0.p, k0sq.data(), sizeof(zd) * slots.size(), s);
            upload(ddz.p, dzs.data(), sizeof(double) * slots.size(), s);
            hipLaunchKernelGGL(k_disk_diag, dim3((unsigned)slots.size()), dim3(256), 0, s, P, dslots.as<int>(),
                               dflag.as<int>(), mc.G.as<zd>(), dk0.as<zd>(), ddz.as<double>(), md.kt.as<zd>(), n,
                               nullptr, nullptr);
            BF_CHECK_LAUNCH();
            BF_HIP(hipStreamSynchronize(s));
        }
        // 1D propagator in air: prop(x) = x * exp(-1im*k0*dz)  ->  P = phase * Gram
        if (!onedim.empty()) {
            std::vector<int> sl;
            std::vector<zd> ph;
            for (int f = 0; f < nb; f++)
                for (int r : onedim) {
                    const zd k0 = k0_of(eps[r], 299792458.0 / fr[f0 + f]);
                    const zd e = neg_i(k0);
                    sl.push_back((int)((size_t)f * nreg + r));
                    ph.push_back(cexp_j(cmk<double>(distance[r] * e.re, distance[r] * e.im)));
                }
            if (dslots.bytes < sizeof(int) * sl.size()) dslots.alloc(sizeof(int) * sl.size());
            if (dk0.bytes < sizeof(zd) * sl.size()) dk0.alloc(sizeof(zd) * sl.size());
            upload(dslots.p, sl.data(), sizeof(int) * sl.size(), s);
            upload(dk0.p, ph.data(), sizeof(zd) * sl.size(), s);
            hipLaunchKernelGGL(k_scaled_copy, dim3((unsigned)sl.size()), dim3(256), 0, s, P, dslots.as<int>(),
                               dk0.as<zd>(), mc.G.as<zd>(), (int)nn);
            BF_CHECK_LAUNCH();
            BF_HIP(hipStreamSynchronize(s));
        }
        // explicitly propagated air regions
        if (!explicit_air.empty()) {
            if (prop.bytes < sizeof(zd) * n * Ktot) prop.alloc(sizeof(zd) * n * Ktot);
            if (kind == BF_PROP_MOMENTUM && scr2.bytes < sizeof(zd) * n * Ktot) scr2.alloc(sizeof(zd) * n * Ktot);
            for (int f = 0; f < nb; f++)
                for (int r : explicit_air) {
                    hipLaunchKernelGGL(k_copy_z, dim3(g1(n * Ktot, 4096)), dim3(256), 0, s, prop.as<zd>(),
                                       md.pat.as<zd>(), n * Ktot);
                    BF_CHECK_LAUNCH();
                    propagate_fields(ctx, g, kind, diskR, surfaces_dev, prop.as<zd>(), n * md.C, eps[r], distance[r],
                                     tx[r], ty[r], flat[r] ? -1 : (int64_t)(K * r), 299792458.0 / fr[f0 + f],
                                     scr2.as<zd>());
                    gram.run({GemmProb{md.pat.as<zd>(), prop.as<zd>(), nullptr, (int)((size_t)f * nreg + r), 0}}, P);
                }
        }
        P += (size_t)nb * nreg * nn;
    }
    if (!queue) parseval_flush(ctx, mc, n, 1.0 / (double)((int64_t)g.nx * g.ny), local_q);
    // local buffers are released on return: wait for the work that reads them (with a
    // queue and none of them, the caller's flush and chains follow in stream order)
    if (!queue || Mom.p || ktabs.p || rtabs.p || scratch.p || prop.p || scr2.p) BF_HIP(hipStreamSynchronize(s));
}

// inv(G) of the mode set's Gram matrix, formed on first use; usable for the flat disks' inverses when
// G is well conditioned (1-norm condition number below 100: the products inv(G) diag(1/ph) then carry
// the LU's accuracy; non-orthogonal user patterns with a worse G keep the LU of every disk matrix)
bool ginv_usable(bf_ctx* ctx, ModeCache& mc, const ModesDev& md) {
    if (!mc.haveG || md.n > BF_MAX_MODES) return false;
    if (!mc.haveGinv) {
        const int n = md.n;
        const size_t nn = (size_t)n * n;
        mc.Ginv.alloc(sizeof(zd) * nn);
        launch_inverse(ctx, mc.G.as<zd>(), mc.Ginv.as<zd>(), n, nullptr, 1);
        std::vector<zd> g(nn), gi(nn);
        download(g.data(), mc.G.p, sizeof(zd) * nn, ctx->stream);
        download(gi.data(), mc.Ginv.p, sizeof(zd) * nn, ctx->stream);
        BF_HIP(hipStreamSynchronize(ctx->stream));
        auto norm1 = [&](const std::vector<zd>& a) {
            double mx = 0.0;
            for (int j = 0; j < n; j++) {
                double c = 0.0;
                for (int i = 0; i < n; i++) c += std::hypot(a[(size_t)j * n + i].re, a[(size_t)j * n + i].im);
                mx = std::max(mx, c);
            }
            return mx;
        };
        const double cond = norm1(g) * norm1(gi);
        mc.ginv_ok = std::isfinite(cond) && cond < 100.0;
        mc.haveGinv = true;
    }
    return mc.ginv_ok;
}

void disk_diag_flush(bf_ctx* ctx, ModeCache& mc, const ModesDev& md, ParsevalQueue& q, Stager* st, zd* pinv_base) {
    if (q.dslot.empty()) return;
    hipStream_t s = ctx->stream;
    const size_t m = q.dslot.size();
    const zd* ginv = pinv_base && ginv_usable(ctx, mc, md) ? mc.Ginv.as<zd>() : nullptr;
    DevBuf& dslots = scratch_buf(ctx, SB_SLOTS, sizeof(int) * m);
    DevBuf& dflag = scratch_buf(ctx, SB_FLAG, sizeof(int) * m);
    DevBuf& dk0 = scratch_buf(ctx, SB_K0, sizeof(zd) * m);
    DevBuf& ddz = scratch_buf(ctx, SB_DZ, sizeof(double) * m);
    stage_upload(st, dslots.p, q.dslot.data(), sizeof(int) * m, s);
    stage_upload(st, dflag.p, q.dfrom.data(), sizeof(int) * m, s);
    stage_upload(st, dk0.p, q.dk0sq.data(), sizeof(zd) * m, s);
    stage_upload(st, ddz.p, q.ddz.data(), sizeof(double) * m, s);
    hipLaunchKernelGGL(k_disk_diag, dim3((unsigned)m), dim3(256), 0, s, q.diag_base, dslots.as<int>(), dflag.as<int>(),
                       mc.G.as<zd>(), dk0.as<zd>(), ddz.as<double>(), md.kt.as<zd>(), md.n, ginv ? pinv_base : nullptr,
                       ginv);
    BF_CHECK_LAUNCH();
    if (!st) BF_HIP(hipStreamSynchronize(s));  // the queue's host arrays are cleared next
    q.dslot.clear();
    q.dfrom.clear();
    q.dk0sq.clear();
    q.ddz.clear();
}

// Splits of the group sum: blocks (row blocks x column blocks x splits) fill the
// launch's last round of resident blocks to at least 90% where a few splits can
// (the machine holds 256 CUs x 2 blocks of the 128-entry shape; the 256-entry
// shape, half the blocks at one per CU, fills the same rounds); partial sums go
// to a workspace and k_pgemm_reduce adds them in split order.  The split does
// not depend on the shape, so both shapes sum every entry in the same order.
// Returns the split count and the chunks per split.
static std::pair<int, int> parseval_splits(size_t nmb, size_t ncb, size_t NG, int kc) {
    const int nch = (int)((NG + kc - 1) / kc);
    const double slots = 512.0;
    int best = 1;
    double best_eff = 0.0;
    for (int sp = 1; sp <= 8 && sp <= nch / 16; sp++) {
        const double rounds = (double)(nmb * ncb * sp) / slots;
        const double eff = rounds / std::ceil(rounds);
        if (eff > best_eff + 1e-9) { best = sp; best_eff = eff; }
        if (eff >= 0.9) break;
    }
    const int cps = (nch + best - 1) / best;
    return {(nch + cps - 1) / cps, cps};
}

// the product over the packed upper triangles when the mode set's Grams are Hermitian (option pgemm_herm)
static bool parseval_herm(const ModeCache& mc) { return mc.herm && opts().pgemm_herm; }
// GEMM columns of the product: n (n + 1) / 2 packed entries, or all n^2
static size_t parseval_cols(const ModeCache& mc, size_t nn) { return parseval_herm(mc) ? (size_t)mc.nt : nn; }

template <bool TAB>
static void launch_parseval_gemm(bf_ctx* ctx, int wn, hipStream_t s, const ModeCache& mc, size_t nn, size_t NG,
                                 const PwDesc* pd, const int* slot, const int* nk, const zd* k0, size_t M,
                                 double scale, zd* const* dst, int nmb, const zd* wtab) {
    const bool herm = parseval_herm(mc);
    const size_t ncol = parseval_cols(mc, nn);
    const size_t BN = herm ? (wn == 2 ? PgShape<2, true>::BN : PgShape<4, true>::BN)
                           : (wn == 2 ? PgShape<2>::BN : PgShape<4>::BN);
    const int kc = herm ? PgShape<2, true, TAB>::KC : PgShape<2, false, TAB>::KC;
    const int ncb = (int)((ncol + BN - 1) / BN);
    // the group split follows from the narrower width's column blocks whatever wn is, so that both widths
    // sum every entry in the same order (bitwise the same results)
    const size_t bn2 = herm ? PgShape<2, true>::BN : PgShape<2>::BN;
    const auto sp = parseval_splits((size_t)nmb, (ncol + bn2 - 1) / bn2, NG, kc);
    zd* part = nullptr;
    if (sp.first > 1) part = scratch_buf(ctx, SB_PGPART, sizeof(zd) * (size_t)sp.first * M * nn).as<zd>();
    const dim3 grid((unsigned)((size_t)nmb * ncb * sp.first));
    const zd* Hs = herm ? mc.Hp.as<zd>() : mc.H.as<zd>();
    const int n = (int)std::lround(std::sqrt((double)nn));
    const int* tri = herm ? mc.tri.as<int>() : nullptr;
#define BF_PGEMM(W, HM)                                                                                         \
    hipLaunchKernelGGL((k_parseval_gemm<W, TAB, HM>), grid, dim3(PgShape<W, HM, TAB>::NT), 0, s, Hs, (int)nn, \
                       (int)NG, \
                       mc.gkx.as<double>(), mc.gky.as<double>(), pd, slot, nk, k0, (int)M, scale, dst, nmb, wtab,  \
                       ncb, sp.second, part, (int)ncol, n, tri)
    if (herm) {
        if (wn == 2)
            BF_PGEMM(2, true);
        else
            BF_PGEMM(4, true);
    } else {
        if (wn == 2)
            BF_PGEMM(2, false);
        else
            BF_PGEMM(4, false);
    }
#undef BF_PGEMM
    BF_CHECK_LAUNCH();
    if (part) {
        hipLaunchKernelGGL(k_pgemm_reduce, dim3((unsigned)((nn + 255) / 256), (unsigned)M), dim3(256), 0, s, part,
                           sp.first, (int)M, (int)nn, dst);
        BF_CHECK_LAUNCH();
    }
}

// Column-block width of the Parseval product for n x n matrices: the tuned one
// if the autotuner has timed this size, else 2 (128 entries per block) unless
// 256 fill the matrix better
static int parseval_wn(bf_ctx* ctx, const ModeCache& mc, size_t nn, size_t NG) {
    int wn = opts().pg_wn;
    const std::string key = "pgemm nn" + std::to_string(nn) + " G" + std::to_string(NG);
    if (wn != 2 && wn != 4 && ctx->tuned.count(key)) wn = (int)ctx->tuned[key][0];
    if (wn != 2 && wn != 4) {
        const size_t c = parseval_cols(mc, nn), b2 = parseval_herm(mc) ? PgShape<2, true>::BN : PgShape<2>::BN;
        wn = ((c + b2 - 1) / b2 * b2 <= (c + 2 * b2 - 1) / (2 * b2) * (2 * b2)) ? 2 : 4;
    }
    return wn;
}

// ---------------------------------------------------------------------------
// Air gaps by a series in the spacing.  For one k0 the weight of group g,
// dz -> exp(-1im*kz_g*dz), is an entire function of the spacing.  Around the
// centre c of a spacing bin [c - h, c + h], with t = (dz - c)/h and
// beta_g = -1im*kz_g*h,
//     exp(-1im*kz_g*dz) = exp(-1im*kz_g*c) * sum_p (beta_g t)^p / p!,
// so P(dz) = sum_p t^p D_p with D_p = sum_g exp(-1im*kz_g*c) beta_g^p / p! * scale * H_g:
// Q coefficient matrices per (k0^2, bin), formed by the same product on a
// table of weights, stand in for one product per gap, and each gap's matrix is
// a Horner sum of Q terms.  The bins have a fixed width per run (2/kmax for up
// to twice the first flush's k0, so |beta_g| <= 1 for every group,
// |kz|^2 <= |k0^2| + max(kx^2 + ky^2)), so every flush of a run reuses the
// coefficients of the bins it shares with earlier flushes; Q is the first
// count whose remainder bound e^|beta| |beta|^Q / Q! is below 1e-17 (the
// weights are at most 1 in modulus for k0^2 in the lower half plane and
// dz >= 0).  dz = 0 is its own bin (one term, the same for every k0: the
// weight is exactly the scale).  The matrices agree with the direct product to
// rounding (a few ulp per weight instead of one).
// ---------------------------------------------------------------------------
struct SeriesKey {
    zd k0sq;
    double c, h;  // centre and half width of the bin
    int q, row0;  // terms, first coefficient row in the table being formed
    const zd* d;  // coefficient matrices (q of them, nn apart)
    const zd* e;  // coefficient matrices of the inverse (q of them), or null
};
struct SeriesProb {
    zd* dst;
    double t;  // (dz - c)/h
    int key, pad;
};
struct SeriesSlice {
    int key, first, count, pad;
};
constexpr int SERIES_QMAX = 24;
constexpr int SERIES_SLICE = 64;

// weight table of the coefficient rows, wt[g][row0 + p] = exp(-1im*kz*c) * beta^p / p! * scale
__global__ void __launch_bounds__(256) k_series_coef(const SeriesKey* __restrict__ keys,
                                                     const double* __restrict__ gkx, const double* __restrict__ gky,
                                                     int NG, int nrows, double scale, zd* __restrict__ wt) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= NG) return;
    const SeriesKey k = keys[blockIdx.y];
    const zd kz = kz_of(k.k0sq, gkx[g], gky[g]);
    const zd e = kz_phase_of(kz, k.c);
    const zd beta = cmk<double>(kz.im * k.h, -kz.re * k.h);
    zd b = cmk<double>(e.re * scale, e.im * scale);
    zd* out = wt + (size_t)g * nrows + k.row0;
    out[0] = b;
    for (int p = 1; p < k.q; p++) {
        b = cmul(b, beta);
        b = cmk<double>(b.re / p, b.im / p);
        out[p] = b;
    }
}

// each gap's matrix from its bin's coefficient matrices, Horner in t;
// block: 128 entries of the slice's problems (all of one key)
__global__ void __launch_bounds__(128) k_series_eval(const SeriesSlice* __restrict__ sl,
                                                     const SeriesProb* __restrict__ pr,
                                                     const SeriesKey* __restrict__ keys, int nn) {
    const int e = blockIdx.x * 128 + threadIdx.x;
    if (e >= nn) return;
    const SeriesSlice S = sl[blockIdx.y];
    const SeriesKey k = keys[S.key];
    zd d[SERIES_QMAX];
#pragma unroll
    for (int p = 0; p < SERIES_QMAX; p++) d[p] = p < k.q ? k.d[(size_t)p * nn + e] : cmk<double>(0.0, 0.0);
    for (int i = 0; i < S.count; i++) {
        const SeriesProb P = pr[S.first + i];
        zd acc = cmk<double>(0.0, 0.0);  // terms beyond q are zero and leave acc zero
#pragma unroll
        for (int p = SERIES_QMAX - 1; p >= 0; p--)
            acc = cmk<double>(d[p].re + P.t * acc.re, d[p].im + P.t * acc.im);
        P.dst[e] = acc;
    }
}

// The inverse of a gap's matrix by the same kind of series.  With P(c + t h) = sum_p t^p D_p, the
// coefficients of t^p in P(c + t h) inv(P(c + t h)) = I give inv(P(c + t h)) = sum_p t^p E_p with
//     E_0 = inv(D_0),    E_p = -E_0 sum_{j=1..p} D_j E_{p-j},
// q terms like P's own series.  P is entire in the spacing and invertible on and near the real axis,
// so the E_p fall off like the D_p (|beta|^p / p!); each block checks that the last two terms of its
// columns are below SERIES_INV_TAIL of E_0's (largest |re| + |im|) and otherwise raises *bad, so that
// the caller takes the LU for the run instead.  Column l of every E_p needs only the D_j and column l
// of the earlier E's, so block (key, y) computes columns SINV_CW y .. SINV_CW (y + 1) - 1 on its own:
// E_0 in LDS, D_j and the block's columns of E_{p-j} staged through LDS for each (p, j); every entry
// sums over j, then k, in increasing order.
constexpr double SERIES_INV_TAIL = 1e-14;
constexpr int SINV_CW = 5;  // columns per block
struct SeriesInvJob {
    const zd* d;  // D_0 .. D_{q-1}
    zd* e;        // E_0 (already inverted) .. E_{q-1}
    int q, pad;
};
constexpr int SINV_EPT = (INV_WAVE_MAX * SINV_CW + 127) / 128;  // entries per thread

__global__ void __launch_bounds__(128) k_series_inverse(const SeriesInvJob* __restrict__ jobs, int n, int* bad) {
    extern __shared__ unsigned char smem_raw[];
    const int nn = n * n;
    const int c0 = blockIdx.y * SINV_CW, cw = min(SINV_CW, n - c0), ne = n * cw;
    zd* E0 = reinterpret_cast<zd*>(smem_raw);  // column-major n x n
    zd* A = E0 + nn;                           // D_j
    zd* Bm = A + nn;                           // the block's columns of E_{p-j}, then of the sum S (n x cw)
    __shared__ double red[128];
    const SeriesInvJob J = jobs[blockIdx.x];
    const int tid = threadIdx.x;
    double m0 = 0.0, mt = 0.0;  // largest |re| + |im| in the block's columns of E_0 and of the last two terms
    for (int e = tid; e < nn; e += 128) {
        const zd v = J.e[e];
        E0[e] = v;
        const int l = e / n;
        if (l >= c0 && l < c0 + cw) m0 = fmax(m0, fabs(v.re) + fabs(v.im));
    }
    for (int p = 1; p < J.q; p++) {
        zd acc[SINV_EPT];
#pragma unroll
        for (int r = 0; r < SINV_EPT; r++) acc[r] = cmk<double>(0.0, 0.0);
        for (int j = 1; j <= p; j++) {
            __syncthreads();
            const zd* Dj = J.d + (size_t)j * nn;
            const zd* Ep = J.e + (size_t)(p - j) * nn + (size_t)c0 * n;  // columns c0 .. c0 + cw - 1
            for (int e = tid; e < nn; e += 128) A[e] = Dj[e];
            for (int e = tid; e < ne; e += 128) Bm[e] = Ep[e];
            __syncthreads();
#pragma unroll
            for (int r = 0; r < SINV_EPT; r++) {
                const int e = tid + 128 * r;
                if (e < ne) {
                    const int i = e % n, l = e / n;
                    zd a = acc[r];
                    for (int k = 0; k < n; k++) a = cmac(a, A[i + n * k], Bm[k + n * l]);
                    acc[r] = a;
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < SINV_EPT; r++) {
            const int e = tid + 128 * r;
            if (e < ne) Bm[e] = acc[r];
        }
        __syncthreads();
        zd* Eo = J.e + (size_t)p * nn + (size_t)c0 * n;
#pragma unroll
        for (int r = 0; r < SINV_EPT; r++) {
            const int e = tid + 128 * r;
            if (e < ne) {
                const int i = e % n, l = e / n;
                zd a = cmk<double>(0.0, 0.0);
                for (int k = 0; k < n; k++) a = cmac(a, E0[i + n * k], Bm[k + n * l]);
                a = cmk<double>(-a.re, -a.im);
                Eo[e] = a;
                if (p >= J.q - 2) mt = fmax(mt, fabs(a.re) + fabs(a.im));
            }
        }
        // the block reads its E_p back for the next terms: the stores above are visible to the block
        // after the barrier at the top of the next j loop
    }
    red[tid] = m0;
    __syncthreads();
    for (int o = 64; o > 0; o >>= 1) {
        if (tid < o) red[tid] = fmax(red[tid], red[tid + o]);
        __syncthreads();
    }
    const double e0max = red[0];
    __syncthreads();
    red[tid] = mt;
    __syncthreads();
    for (int o = 64; o > 0; o >>= 1) {
        if (tid < o) red[tid] = fmax(red[tid], red[tid + o]);
        __syncthreads();
    }
    if (tid == 0 && !(red[0] <= SERIES_INV_TAIL * e0max)) *bad = 1;
}

// k_series_inverse on the matrix cores for 8 < n <= 32 (operands padded to 32; KF = ceil(n / 4) k steps):
// block (key, column block y) computes columns 16 y .. 16 y + 15 of every E_p, wave w rows 16 w .. 16 w + 15
// (for n <= 16 wave 1's rows are all beyond n: its products are clamped copies and never stored).
// S_p = sum_j D_j E_{p-j} and E_p = -E_0 S_p by v_mfma_f64_16x16x4f64, three MFMAs per complex 4-k step;
// the lanes load D_j and E_{p-j} (global memory, L2) in the MFMA layouts with indices clamped into the
// matrix (B entries at k >= n zeroed), the next j's while the current j's products run; S_p goes through
// LDS to the second product.  The same tail check as k_series_inverse.
template <int KF>
__global__ void __launch_bounds__(128) k_series_inverse_mfma(const SeriesInvJob* __restrict__ jobs, int n,
                                                             int* bad) {
    constexpr int LD = 33;
    __shared__ zd S[32 * LD];  // S_p[k][c] at S[k + LD c] (rows k >= n stay zero)
    __shared__ double red[128];
    const SeriesInvJob J = jobs[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, kq = lane >> 4, c = lane & 15;
    const int c0 = blockIdx.y * 16, nn = n * n;
    const int ar = 16 * w + c, bcol = c0 + c;  // A row of the lane, B / output column
    const int rr = ar < n ? ar : n - 1, cc = bcol < n ? bcol : n - 1;
    for (int e = tid; e < 32 * LD; e += 128) S[e] = cmk<double>(0.0, 0.0);
    // E_0's A operand (rows ar, k = 4 s + kq) stays in registers
    zd e0a[KF];
#pragma unroll
    for (int s = 0; s < KF; s++) {
        const int k = 4 * s + kq;
        e0a[s] = J.e[rr + n * (k < n ? k : n - 1)];
    }
    double m0 = 0.0, mt = 0.0;
    const int cw = min(16, n - c0);
    for (int e = tid; e < n * cw; e += 128) {
        const zd v = J.e[e % n + (size_t)n * (c0 + e / n)];
        m0 = fmax(m0, fabs(v.re) + fabs(v.im));
    }
    auto gauss = [](d4& cr, d4& ci, d4& cg, const zd& a, const zd& b) {
        cr = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, b.re, cr, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f64_16x16x4f64(a.im, b.im, ci, 0, 0, 0);
        cg = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re + a.im, b.re + b.im, cg, 0, 0, 0);
    };
    for (int p = 1; p < J.q; p++) {
        __syncthreads();  // E_{p-1} of both waves written; S of term p - 1 read
        d4 cr = {0.0, 0.0, 0.0, 0.0}, ci = cr, cg = cr;
        zd an[KF], bn[KF];
        auto load = [&](int j) {
            const zd* Dj = J.d + (size_t)j * nn;
            const zd* Ep = J.e + (size_t)(p - j) * nn;
#pragma unroll
            for (int s = 0; s < KF; s++) {
                const int k = 4 * s + kq, kk = k < n ? k : n - 1;
                an[s] = Dj[rr + n * kk];
                bn[s] = Ep[kk + n * cc];
            }
        };
        load(1);
        for (int j = 1; j <= p; j++) {
            zd a[KF], b[KF];
#pragma unroll
            for (int s = 0; s < KF; s++) {
                a[s] = an[s];
                b[s] = bn[s];
            }
            if (j < p) load(j + 1);
#pragma unroll
            for (int s = 0; s < KF; s++) {
                const bool kin = 4 * s + kq < n;
                const zd bz = cmk<double>(kin ? b[s].re : 0.0, kin ? b[s].im : 0.0);
                gauss(cr, ci, cg, a[s], bz);
            }
        }
        // S_p tile: register v holds S[16 w + 4 v + kq][c]
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int i = 16 * w + 4 * v + kq;
            if (i < n) S[i + LD * c] = cmk<double>(cr[v] - ci[v], cg[v] - cr[v] - ci[v]);
        }
        __syncthreads();
        d4 er = {0.0, 0.0, 0.0, 0.0}, ei = er, eg = er;
#pragma unroll
        for (int s = 0; s < KF; s++) gauss(er, ei, eg, e0a[s], S[4 * s + kq + LD * c]);
        zd* Eo = J.e + (size_t)p * nn;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int i = 16 * w + 4 * v + kq;
            if (i < n && bcol < n) {
                const zd a = cmk<double>(-(er[v] - ei[v]), -(eg[v] - er[v] - ei[v]));
                Eo[i + (size_t)n * bcol] = a;
                if (p >= J.q - 2) mt = fmax(mt, fabs(a.re) + fabs(a.im));
            }
        }
    }
    red[tid] = m0;
    __syncthreads();
    for (int o = 64; o > 0; o >>= 1) {
        if (tid < o) red[tid] = fmax(red[tid], red[tid + o]);
        __syncthreads();
    }
    const double e0max = red[0];
    __syncthreads();
    red[tid] = mt;
    __syncthreads();
    for (int o = 64; o > 0; o >>= 1) {
        if (tid < o) red[tid] = fmax(red[tid], red[tid + o]);
        __syncthreads();
    }
    if (tid == 0 && !(red[0] <= SERIES_INV_TAIL * e0max)) *bad = 1;
}

using SeriesKeyId = std::array<uint64_t, 3>;  // k0^2 bits (re, im) and spacing bin; {0, 0, ~0}: dz = 0
constexpr uint64_t SERIES_ZERO_BIN = ~(uint64_t)0;

// The series of key `key` (k0^2 = k0, bin of width mc.series_w): centre, half width and terms; false if its
// k0 is beyond the bins' allowance or if `cnt` gaps do not pay for the terms (at least q + 4 gaps)
static bool series_key_for(const ModeCache& mc, const SeriesKeyId& key, zd k0, int cnt, SeriesKey& out) {
    const double w = mc.series_w;
    const bool zero = key[2] == SERIES_ZERO_BIN;
    const double h = zero ? 0.0 : 0.5 * w, c = zero ? 0.0 : ((double)(int64_t)key[2] + 0.5) * w;
    const double beta = std::sqrt(std::hypot(k0.re, k0.im) + mc.gk2max) * h;
    int Q = 1;
    if (h > 0.0) {
        double term = beta;  // beta^Q / Q!
        while (std::exp(beta) * term >= 1e-17 && Q < SERIES_QMAX) {
            Q++;
            term *= beta / Q;
        }
        if (std::exp(beta) * term >= 1e-17) return false;
    }
    if (cnt < Q + 4) return false;
    out = SeriesKey{k0, c, h, Q, 0, nullptr};
    return true;
}

// The coefficient matrices of the keys `fresh` (rows row0 .. row0 + q - 1 of nrows): weight table, one
// product on it, entries added to mc.series for the rest of the run.  With `bad` (a device flag), also
// the inverse's coefficients (k_series_inverse, which raises the flag if they do not converge).
static void series_form(bf_ctx* ctx, ModeCache& mc, int n, double scale, const std::vector<SeriesKey>& fresh,
                        const std::vector<SeriesKeyId>& fresh_key, int nrows, Stager* st, int* bad = nullptr) {
    if (fresh.empty()) return;
    hipStream_t s = ctx->stream;
    const size_t nn = (size_t)n * n, NG = (size_t)mc.ngroups;
    zd* blk = static_cast<zd*>(mc.series_mem.take(sizeof(zd) * nn * (size_t)nrows));
    std::vector<zd*> rowdst((size_t)nrows);
    for (int r = 0; r < nrows; r++) rowdst[(size_t)r] = blk + (size_t)r * nn;
    DevBuf& dkey = scratch_buf(ctx, SB_SKEY, sizeof(SeriesKey) * fresh.size());
    DevBuf& drow = scratch_buf(ctx, SB_SROWDST, sizeof(zd*) * (size_t)nrows);
    DevBuf& dw = scratch_buf(ctx, SB_SW, sizeof(zd) * NG * (size_t)nrows);
    stage_upload(st, dkey.p, fresh.data(), sizeof(SeriesKey) * fresh.size(), s);
    stage_upload(st, drow.p, rowdst.data(), sizeof(zd*) * rowdst.size(), s);
    {
        TimedLaunch tl(ctx, "pseries", 0.0);
        hipLaunchKernelGGL(k_series_coef, dim3((unsigned)((NG + 255) / 256), (unsigned)fresh.size()), dim3(256), 0, s,
                           dkey.as<SeriesKey>(), mc.gkx.as<double>(), mc.gky.as<double>(), (int)NG, nrows, scale,
                           dw.as<zd>());
        BF_CHECK_LAUNCH();
    }
    {
        TimedLaunch tl(ctx, "pgemm", 8.0 * (double)parseval_cols(mc, nn) * (double)NG * (double)nrows);
        const int nmb = (nrows + PG_BM - 1) / PG_BM;
        launch_parseval_gemm<true>(ctx, parseval_wn(ctx, mc, nn, NG), s, mc, nn, NG, nullptr, nullptr, nullptr, nullptr,
                                   (size_t)nrows, scale, drow.as<zd*>(), nmb, dw.as<zd>());
    }
    zd* inv = nullptr;
    std::vector<int> rows0(fresh.size());
    std::vector<SeriesInvJob> ij(fresh.size());
    if (bad) {
        // E_0 = inv(D_0) by the wave inverse, then the further terms per key
        inv = static_cast<zd*>(mc.series_mem.take(sizeof(zd) * nn * (size_t)nrows));
        int qmax = 1;
        for (size_t k = 0; k < fresh.size(); k++) {
            rows0[k] = fresh[k].row0;
            ij[k] = SeriesInvJob{blk + (size_t)fresh[k].row0 * nn, inv + (size_t)fresh[k].row0 * nn, fresh[k].q, 0};
            qmax = std::max(qmax, fresh[k].q);
        }
        DevBuf& dsel = scratch_buf(ctx, SB_SINVSEL, sizeof(int) * rows0.size());
        DevBuf& dj = scratch_buf(ctx, SB_SINVJ, sizeof(SeriesInvJob) * ij.size());
        stage_upload(st, dsel.p, rows0.data(), sizeof(int) * rows0.size(), s);
        stage_upload(st, dj.p, ij.data(), sizeof(SeriesInvJob) * ij.size(), s);
        TimedLaunch tl(ctx, "sinverse", 8.0 * (double)nn * n * (double)fresh.size() * (1.0 + qmax * (qmax + 1) / 2));
        const size_t smem1 = sizeof(zd) * nn;
        allow_lds((const void*)k_batched_inverse_wave, smem1);
        hipLaunchKernelGGL(k_batched_inverse_wave, dim3((unsigned)fresh.size()), dim3(64), smem1, s, blk, inv, n,
                           dsel.as<int>(), (const InvSel*)nullptr);
        BF_CHECK_LAUNCH();
        if (n > 8 && n <= 32 && opts().sinv_mfma) {
            const dim3 grid((unsigned)fresh.size(), (unsigned)((n + 15) / 16));
            switch ((n + 3) / 4) {
                case 3: hipLaunchKernelGGL(k_series_inverse_mfma<3>, grid, dim3(128), 0, s, dj.as<SeriesInvJob>(), n, bad); break;
                case 4: hipLaunchKernelGGL(k_series_inverse_mfma<4>, grid, dim3(128), 0, s, dj.as<SeriesInvJob>(), n, bad); break;
                case 5: hipLaunchKernelGGL(k_series_inverse_mfma<5>, grid, dim3(128), 0, s, dj.as<SeriesInvJob>(), n, bad); break;
                case 6: hipLaunchKernelGGL(k_series_inverse_mfma<6>, grid, dim3(128), 0, s, dj.as<SeriesInvJob>(), n, bad); break;
                case 7: hipLaunchKernelGGL(k_series_inverse_mfma<7>, grid, dim3(128), 0, s, dj.as<SeriesInvJob>(), n, bad); break;
                default: hipLaunchKernelGGL(k_series_inverse_mfma<8>, grid, dim3(128), 0, s, dj.as<SeriesInvJob>(), n, bad); break;
            }
        } else {
            const size_t smem3 = sizeof(zd) * (2 * nn + (size_t)n * SINV_CW);
            allow_lds((const void*)k_series_inverse, smem3);
            hipLaunchKernelGGL(k_series_inverse, dim3((unsigned)fresh.size(), (unsigned)((n + SINV_CW - 1) / SINV_CW)),
                               dim3(128), smem3, s, dj.as<SeriesInvJob>(), n, bad);
        }
        BF_CHECK_LAUNCH();
    }
    for (size_t k = 0; k < fresh.size(); k++)
        mc.series[fresh_key[k]] = ModeCache::SeriesEntry{blk + (size_t)fresh[k].row0 * nn, fresh[k].q, fresh[k].c,
                                                         fresh[k].h, bad ? inv + (size_t)fresh[k].row0 * nn : nullptr};
    if (!st) BF_HIP(hipStreamSynchronize(s));  // the host arrays go out of scope
}

// Serves the problems of q whose (k0^2, spacing bin) has coefficients from an
// earlier flush of the run, or enough gaps in this flush to pay for them (at
// least q + 4), and leaves the others in q.
static void parseval_series(bf_ctx* ctx, ModeCache& mc, int n, double scale, ParsevalQueue& q, Stager* st) {
    const size_t M = q.pd.size(), nn = (size_t)n * n;
    hipStream_t s = ctx->stream;
    using Key = SeriesKeyId;
    auto bits = [](double v) {
        uint64_t b;
        std::memcpy(&b, &v, 8);
        return b;
    };
    if (mc.series_w <= 0.0) {
        double k2 = 0.0;
        for (size_t m = 0; m < M; m++) k2 = std::max(k2, std::hypot(q.pd[m].k0sq.re, q.pd[m].k0sq.im));
        const double kb = std::sqrt(4.0 * k2 + mc.gk2max);
        if (!(kb > 0.0) || !std::isfinite(kb)) return;
        mc.series_w = 2.0 / kb;
    }
    const double w = mc.series_w;
    // every problem's key as a dense index: 0 for dz = 0, else 1 + k0 slot * nb + (bin - bmin), where
    // consecutive problems mostly share their k0 (a geometry's gaps at one frequency); -1: direct
    std::vector<int> slot(M, -1);
    std::vector<int64_t> bin(M, 0);
    std::map<std::pair<uint64_t, uint64_t>, int> k0map;
    std::vector<zd> k0v;
    int last = -1;
    uint64_t lre = 0, lim = 0;
    int64_t bmin = INT64_MAX, bmax = INT64_MIN;
    bool any_zero = false;
    for (size_t m = 0; m < M; m++) {
        const zd k0 = q.pd[m].k0sq;
        const double dz = q.pd[m].dz;
        if (!(k0.im <= 0.0) || !std::isfinite(k0.re)) continue;
        if (dz == 0.0) {
            slot[m] = -2;
            any_zero = true;
            continue;
        }
        if (!(dz > 0.0) || !std::isfinite(dz) || dz / w > 1e15) continue;
        const uint64_t re = bits(k0.re), im = bits(k0.im);
        if (last < 0 || re != lre || im != lim) {
            auto it = k0map.find({re, im});
            if (it == k0map.end()) {
                it = k0map.emplace(std::make_pair(re, im), (int)k0v.size()).first;
                k0v.push_back(k0);
            }
            last = it->second;
            lre = re;
            lim = im;
        }
        slot[m] = last;
        bin[m] = (int64_t)std::floor(dz / w);
        bmin = std::min(bmin, bin[m]);
        bmax = std::max(bmax, bin[m]);
    }
    const int64_t nb = bmax >= bmin ? bmax - bmin + 1 : 0;
    const bool dense_ok = nb > 0 && nb <= 4096 && (int64_t)k0v.size() * nb <= ((int64_t)1 << 22);
    const size_t nidx = 1 + (dense_ok ? k0v.size() * (size_t)nb : 0);
    auto idx_of = [&](size_t m) -> int64_t {
        if (slot[m] == -2) return 0;
        if (slot[m] < 0 || !dense_ok) return -1;
        return 1 + (int64_t)slot[m] * nb + (bin[m] - bmin);
    };
    std::vector<int> cnt(nidx, 0);
    for (size_t m = 0; m < M; m++) {
        const int64_t i = idx_of(m);
        if (i >= 0) cnt[(size_t)i]++;
    }
    auto key_at = [&](size_t i, zd& k0) -> Key {
        if (i == 0) {
            k0 = k0v.empty() ? cmk<double>(0.0, -0.0) : k0v[0];
            for (size_t m = 0; m < M; m++)
                if (slot[m] == -2) {
                    k0 = q.pd[m].k0sq;
                    break;
                }
            return {0, 0, SERIES_ZERO_BIN};
        }
        const size_t sl = (i - 1) / (size_t)nb;
        k0 = k0v[sl];
        return {bits(k0.re), bits(k0.im), (uint64_t)(bmin + (int64_t)((i - 1) % (size_t)nb))};
    };
    (void)any_zero;
    // new keys worth forming
    std::vector<SeriesKey> fresh;
    std::vector<SeriesKeyId> fresh_key;
    int nrows = 0;
    for (size_t i = 0; i < nidx; i++) {
        if (!cnt[i]) continue;
        zd k0;
        const Key key = key_at(i, k0);
        if (mc.series.count(key)) continue;
        SeriesKey sk;
        if (!series_key_for(mc, key, k0, cnt[i], sk)) continue;
        sk.row0 = nrows;
        fresh.push_back(sk);
        fresh_key.push_back(key);
        nrows += sk.q;
    }
    series_form(ctx, mc, n, scale, fresh, fresh_key, nrows, st);
    // problems with coefficients: ordered by key (counting sort), slices of SERIES_SLICE; the rest stays in q
    std::vector<int> key_of_idx(nidx, -1);
    std::vector<SeriesKey> keys;
    for (size_t i = 0; i < nidx; i++) {
        if (!cnt[i]) continue;
        zd k0;
        const Key key = key_at(i, k0);
        auto c = mc.series.find(key);
        if (c == mc.series.end()) continue;
        key_of_idx[i] = (int)keys.size();
        keys.push_back(SeriesKey{k0, c->second.c, c->second.h, c->second.q, 0, c->second.d});
    }
    std::vector<int> pkey(M, -1);
    for (size_t m = 0; m < M; m++) {
        const int64_t i = idx_of(m);
        if (i >= 0) pkey[m] = key_of_idx[(size_t)i];
    }
    if (keys.empty()) return;
    std::vector<int> kfirst(keys.size() + 1, 0);
    for (size_t m = 0; m < M; m++)
        if (pkey[m] >= 0) kfirst[(size_t)pkey[m] + 1]++;
    for (size_t k = 0; k < keys.size(); k++) kfirst[k + 1] += kfirst[k];
    std::vector<SeriesProb> probs((size_t)kfirst.back());
    std::vector<int> fill(kfirst.begin(), kfirst.end() - 1);
    size_t keep = 0;
    for (size_t m = 0; m < M; m++) {
        const int k = pkey[m];
        if (k < 0) {
            q.pd[keep] = q.pd[m];
            q.dst[keep] = q.dst[m];
            keep++;
            continue;
        }
        const SeriesKey& K = keys[(size_t)k];
        probs[(size_t)fill[k]++] = SeriesProb{q.dst[m], K.h > 0.0 ? (q.pd[m].dz - K.c) / K.h : 0.0, k, 0};
    }
    q.pd.resize(keep);
    q.dst.resize(keep);
    std::vector<SeriesSlice> slices;
    for (size_t k = 0; k < keys.size(); k++)
        for (int f = kfirst[k]; f < kfirst[k + 1]; f += SERIES_SLICE)
            slices.push_back(SeriesSlice{(int)k, f, std::min(SERIES_SLICE, kfirst[k + 1] - f), 0});
    DevBuf& dkey = scratch_buf(ctx, SB_SD, sizeof(SeriesKey) * keys.size());
    DevBuf& dprob = scratch_buf(ctx, SB_SPROB, sizeof(SeriesProb) * probs.size());
    DevBuf& dsl = scratch_buf(ctx, SB_SSLICE, sizeof(SeriesSlice) * slices.size());
    stage_upload(st, dkey.p, keys.data(), sizeof(SeriesKey) * keys.size(), s);
    stage_upload(st, dprob.p, probs.data(), sizeof(SeriesProb) * probs.size(), s);
    stage_upload(st, dsl.p, slices.data(), sizeof(SeriesSlice) * slices.size(), s);
    {
        TimedLaunch tl(ctx, "pseries", 8.0 * (double)nn * (double)probs.size() * SERIES_QMAX);
        hipLaunchKernelGGL(k_series_eval, dim3((unsigned)((nn + 127) / 128), (unsigned)slices.size()), dim3(128), 0,
                           s, dsl.as<SeriesSlice>(), dprob.as<SeriesProb>(), dkey.as<SeriesKey>(), (int)nn);
        BF_CHECK_LAUNCH();
    }
    if (!st) BF_HIP(hipStreamSynchronize(s));  // the host arrays go out of scope
}

void parseval_flush(bf_ctx* ctx, ModeCache& mc, int n, double scale, ParsevalQueue& q, Stager* st) {
    if (q.pd.empty()) return;
    if (opts().parseval_series) parseval_series(ctx, mc, n, scale, q, st);
    if (q.pd.empty()) return;
    hipStream_t s = ctx->stream;
    const size_t nn = (size_t)n * n, NG = (size_t)mc.ngroups, M = q.pd.size();
    if (M > (size_t)INT32_MAX / 2) throw Error(BF_ERR_UNSUPPORTED, "too many Parseval problems in one flush");
    const int nmb = (int)((M + PG_BM - 1) / PG_BM);
    // the distinct k0^2 of each block of PG_BM problems (bitwise) and every problem's slot among them
    std::vector<int> slot(M), nk(nmb);
    std::vector<zd> k0s((size_t)nmb * PG_BM, cmk<double>(0.0, 0.0));
    for (int b = 0; b < nmb; b++) {
        int cnt = 0;
        zd* list = k0s.data() + (size_t)b * PG_BM;
        for (size_t m = (size_t)b * PG_BM; m < std::min(M, (size_t)(b + 1) * PG_BM); m++) {
            const zd v = q.pd[m].k0sq;
            int k = 0;
            while (k < cnt && !(std::memcmp(&list[k], &v, sizeof(zd)) == 0)) k++;
            if (k == cnt) list[cnt++] = v;
            slot[m] = k;
        }
        nk[b] = cnt;
    }
    DevBuf& dpw = scratch_buf(ctx, SB_PW, sizeof(PwDesc) * M);
    DevBuf& ddst = scratch_buf(ctx, SB_DST, sizeof(zd*) * M);
    DevBuf& dslot = scratch_buf(ctx, SB_PGSLOT, sizeof(int) * M);
    DevBuf& dnk = scratch_buf(ctx, SB_PGNK, sizeof(int) * nmb);
    DevBuf& dk0 = scratch_buf(ctx, SB_PGK0, sizeof(zd) * k0s.size());
    stage_upload(st, dpw.p, q.pd.data(), sizeof(PwDesc) * M, s);
    stage_upload(st, ddst.p, q.dst.data(), sizeof(zd*) * M, s);
    stage_upload(st, dslot.p, slot.data(), sizeof(int) * M, s);
    stage_upload(st, dnk.p, nk.data(), sizeof(int) * nmb, s);
    stage_upload(st, dk0.p, k0s.data(), sizeof(zd) * k0s.size(), s);
    auto launch = [&](int w) {
        launch_parseval_gemm<false>(ctx, w, s, mc, nn, NG, dpw.as<PwDesc>(), dslot.as<int>(), dnk.as<int>(),
                                    dk0.as<zd>(), M, scale, ddst.as<zd*>(), nmb, nullptr);
    };
    // waves along the entries; with the autotuner, both shapes are timed on the first flush of a matrix
    // size with at least 256 blocks of 128 entries (they compute the same sums in the same order, so the second launch rewrites identical
    // values)
    int wn = opts().pg_wn;
    const std::string key = "pgemm nn" + std::to_string(nn) + " G" + std::to_string(NG);
    if (wn != 2 && wn != 4 && !ctx->tuned.count(key) && opts().autotune &&
        (size_t)nmb * ((nn + PgShape<2>::BN - 1) / PgShape<2>::BN) >= 256) {
        const auto t0 = std::chrono::steady_clock::now();
        double best = 1e300;
        for (int cand : {2, 4}) {
            hipEvent_t e0, e1;
            BF_HIP(hipEventCreate(&e0));
            BF_HIP(hipEventCreate(&e1));
            BF_HIP(hipEventRecord(e0, s));
            launch(cand);
            BF_HIP(hipEventRecord(e1, s));
            BF_HIP(hipEventSynchronize(e1));
            float ms = 0.0f;
            BF_HIP(hipEventElapsedTime(&ms, e0, e1));
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            if (ms < best) { best = ms; wn = cand; }
        }
        ctx->tuned[key] = {wn};
        ctx->tune_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        save_tuned(ctx, key);
    }
    wn = parseval_wn(ctx, mc, nn, NG);
    {
        TimedLaunch tl(ctx, "pgemm", 8.0 * (double)parseval_cols(mc, nn) * (double)NG * (double)M);
        launch(wn);
    }
    if (!st) BF_HIP(hipStreamSynchronize(s));  // the host arrays are reused by the next flush's uploads
    q.pd.clear();
    q.dst.clear();
}

void upload_surfaces(bf_ctx* ctx, const bf_setup* st, size_t K, DevBuf& surf, std::vector<bool>& flat) {
    hipStream_t s = ctx->stream;
    flat.assign(st->nreg, true);
    if (st->surfaces) {
        const zd* sf = reinterpret_cast<const zd*>(st->surfaces);
        for (int r = 0; r < st->nreg; r++)
            for (size_t i = 0; i < K; i++)
                if (sf[(size_t)r * K + i].re != 0.0 || sf[(size_t)r * K + i].im != 0.0) {
                    flat[r] = false;
                    break;
                }
        surf.alloc(sizeof(zd) * K * st->nreg);
        upload(surf.p, st->surfaces, sizeof(zd) * K * st->nreg, s);
    }
}

static void setup_common(bf_ctx* ctx, const bf_setup* st, const bf_coords* co, Grid& g, DevBuf& surf,
                         std::vector<bool>& flat) {
    g.upload(co, ctx->stream);
    upload_surfaces(ctx, st, g.K(), surf, flat);
}

void seed_modes_dev(bf_ctx* ctx, const bf_coords* coords, int M, int L, double diskR, double* patterns, double* kt) {
    hipStream_t s = ctx->stream;
    Grid g;
    g.upload(coords, s);
    const size_t K = g.K();
    const int n = M * (2 * L + 1);
    DevBuf dkt(sizeof(zd) * n), dpat(sizeof(zd) * n * K), nrm(sizeof(double) * n);
    hipLaunchKernelGGL(k_mode_kt, dim3((n + 63) / 64), dim3(64), 0, s, dkt.as<zd>(), M, L, diskR);
    hipLaunchKernelGGL(k_mode_patterns, dim3(g1(K, 1024), n), dim3(256), 0, s, dpat.as<zd>(), dkt.as<zd>(), L, g.nx,
                       g.ny, g.X.as<double>(), g.Y.as<double>(), diskR);
    hipLaunchKernelGGL(k_row_norm2, dim3(n), dim3(1024), 0, s, dpat.as<zd>(), K, nrm.as<double>());
    hipLaunchKernelGGL(k_row_divide_sqrt, dim3(g1(K, 512), n), dim3(256), 0, s, dpat.as<zd>(), K, nrm.as<double>());
    BF_CHECK_LAUNCH();
    download(patterns, dpat.p, sizeof(zd) * n * K, s);
    download(kt, dkt.p, sizeof(zd) * n, s);
    BF_HIP(hipStreamSynchronize(s));
}

static void axion_modes(bf_ctx* ctx, const Grid& g, const ModesDev& md, const double* B, double velocity_x,
                        double diskR, double f, zd* coeffs_dev) {
    hipStream_t s = ctx->stream;
    const size_t Ktot = md.Ktot;
    DevBuf dB, b(sizeof(zd) * Ktot), nrm(sizeof(double));
    if (B) {
        dB.alloc(sizeof(zd) * Ktot);
        upload(dB.p, B, sizeof(zd) * Ktot, s);
    }
    const double Ma = 2.0 * M_PI / (299792458.0 / f);
    hipLaunchKernelGGL(k_axion_source, dim3(g1(Ktot, 2048)), dim3(256), 0, s, b.as<zd>(), dB.as<zd>(), md.C, g.nx,
                       g.ny, g.X.as<double>(), g.Y.as<double>(), diskR, velocity_x, Ma);
    hipLaunchKernelGGL(k_row_norm2, dim3(1), dim3(1024), 0, s, b.as<zd>(), Ktot, nrm.as<double>());
    hipLaunchKernelGGL(k_row_divide_sqrt, dim3(g1(Ktot, 512), 1), dim3(256), 0, s, b.as<zd>(), Ktot,
                       nrm.as<double>());
    hipLaunchKernelGGL(k_project, dim3(md.n), dim3(1024), 0, s, md.pat.as<zd>(), b.as<zd>(), Ktot, coeffs_dev);
    BF_CHECK_LAUNCH();
    BF_HIP(hipStreamSynchronize(s));
}

void field2modes_dev(bf_ctx* ctx, const bf_coords* coords, const bf_modes* modes, const double* B, double velocity_x,
                     double diskR, double f, double* coeffs) {
    hipStream_t s = ctx->stream;
    Grid g;
    g.upload(coords, s);
    ModesDev md;
    md.upload(modes, g.K(), s);
    DevBuf c(sizeof(zd) * md.n);
    axion_modes(ctx, g, md, B, velocity_x, diskR, f, c.as<zd>());
    download(coeffs, c.p, sizeof(zd) * md.n, s);
    BF_HIP(hipStreamSynchronize(s));
}

void modes2field_dev(bf_ctx* ctx, const bf_coords* coords, const bf_modes* modes, const double* coeffs,
                     double* field) {
    hipStream_t s = ctx->stream;
    Grid g;
    g.upload(coords, s);
    ModesDev md;
    md.upload(modes, g.K(), s);
    DevBuf c(sizeof(zd) * md.n), out(sizeof(zd) * md.Ktot);
    upload(c.p, coeffs, sizeof(zd) * md.n, s);
    hipLaunchKernelGGL(k_modes2field, dim3(g1(md.Ktot, 2048)), dim3(256), 0, s, md.pat.as<zd>(), c.as<zd>(), md.n,
                       md.Ktot, out.as<zd>());
    BF_CHECK_LAUNCH();
    download(field, out.p, sizeof(zd) * md.Ktot, s);
    BF_HIP(hipStreamSynchronize(s));
}

void propagation_matrices_dev(bf_ctx* ctx, const bf_setup* st, const bf_coords* co, const bf_modes* modes,
                              const double* freqs, int nfreq, int kind, double diskR, int precision, double* out) {
    (void)precision;
    hipStream_t s = ctx->stream;
    Grid g;
    DevBuf surf;
    std::vector<bool> flat;
    setup_common(ctx, st, co, g, surf, flat);
    ModesDev md;
    md.upload(modes, g.K(), s);
    const int nreg = st->nreg;
    const size_t nn = (size_t)md.n * md.n;
    DevBuf P(sizeof(zd) * nn * nreg * nfreq);
    std::vector<int> regs(nreg);
    for (int r = 0; r < nreg; r++) regs[r] = r;
    propagation_matrices_core(ctx, g, md, kind, diskR, nreg, st->distance, reinterpret_cast<const zd*>(st->eps),
                              st->tilt_x, st->tilt_y, surf.as<zd>(), flat, freqs, nfreq, regs, P.as<zd>());
    download(out, P.p, sizeof(zd) * nn * nreg * nfreq, s);
    BF_HIP(hipStreamSynchronize(s));
}

std::vector<zd> eps_fixed(const bf_setup* st) {
    const zd* e = reinterpret_cast<const zd*>(st->eps);
    std::vector<zd> v(e, e + st->nreg);
    for (auto& z : v)
        if (std::isnan(z.re) || std::isnan(z.im)) z = cmk<double>(1e30, 0.0);  // bdry.eps[isnan] .= 1e30 (:289)
    return v;
}

// step q of the chain over regions nreg-1 .. 1 (0-based; eps with NaN already 1e30)
static ChainStep chain_step(const zd* eps, int nreg, int q) {
    const int sidx = nreg - 1 - q;                      // s in the reference loop
    const int ir = nreg - sidx, ir1 = nreg - sidx - 1;  // idx_reg(s)-1, idx_reg(s+1)-1
    const zd n1 = csqrt_j(eps[ir1]), n0 = csqrt_j(eps[ir]);
    const zd one = cmk<double>(1.0, 0.0);
    const zd S = rmul(0.5, csub(cdiv_h(one, cmul(n1, n1)), cdiv_h(one, cmul(n0, n0))));
    // get_boundary_matrix(n_left = n0, n_right = n1): G = 1/(2 nr) [(nr+nl) I, (nr-nl) I; ...]
    const zd sc = cdiv_h(one, rmul(2.0, n1));
    ChainStep c;
    c.S = S;
    c.a = cmul(sc, cadd(n1, n0));
    c.b = cmul(sc, csub(n1, n0));
    c.reg = ir;
    c.disk = 0;
    return c;
}

void append_chain_steps(const std::vector<zd>& eps, int nreg, std::vector<ChainStep>& steps) {
    for (int q = 0; q < nreg - 1; q++) steps.push_back(chain_step(eps.data(), nreg, q));
}

void inverse_and_chains(bf_ctx* ctx, int n, zd* P, zd* Pinv, const std::vector<int>& sel,
                        const std::vector<ChainJob>& jobs, const std::vector<ChainStep>& steps, const zd* e0,
                        const zd* reflect, zd* boost, zd* refl, Stager* st, const ChainSrc* disks) {
    hipStream_t s = ctx->stream;
    const size_t nn = (size_t)n * n;
    const int N2 = 2 * n;
    DevBuf& dsel = scratch_buf(ctx, SB_SEL, sizeof(int) * sel.size());
    stage_upload(st, dsel.p, sel.data(), sizeof(int) * sel.size(), s);
    if (!sel.empty()) {
        TimedLaunch tl(ctx, "inverse", (8.0 / 3.0 + 8.0) * (double)n * n * n * (double)sel.size());
        launch_inverse(ctx, P, Pinv, n, dsel.as<int>(), sel.size());
    }
    const size_t nj = jobs.size();
    const bool lds = n <= CHAIN_LDS_MAX;
    // Blocks start roughly in index order, so a batch of chains of unequal length (geometries with
    // different disk counts) is laid out longest first: the last blocks to start are then the short
    // ones, and the launch does not end on a few long chains started late.  Each job writes its own
    // outputs, so the order changes nothing else.
    std::vector<ChainJob> ordered;
    const ChainJob* jl = jobs.data();
    if (opts().chain_order && nj > 1 &&
        std::any_of(jobs.begin(), jobs.end(), [&](const ChainJob& j) { return j.nsteps != jobs[0].nsteps; })) {
        ordered = jobs;
        std::stable_sort(ordered.begin(), ordered.end(),
                         [](const ChainJob& x, const ChainJob& y) { return x.nsteps > y.nsteps; });
        jl = ordered.data();
    }
    DevBuf& djobs = scratch_buf(ctx, SB_JOBS, sizeof(ChainJob) * nj);
    DevBuf& dsteps = scratch_buf(ctx, SB_STEPS, sizeof(ChainStep) * steps.size());
    DevBuf& Tb = scratch_buf(ctx, SB_TB, lds ? 0 : sizeof(zd) * 2 * (size_t)N2 * N2 * nj);
    DevBuf& cw = scratch_buf(ctx, SB_CW, lds ? 0 : sizeof(zd) * (nn + 3 * (size_t)n) * nj);
    stage_upload(st, djobs.p, jl, sizeof(ChainJob) * nj, s);
    stage_upload(st, dsteps.p, steps.data(), sizeof(ChainStep) * steps.size(), s);
    double flops = 0.0;
    for (const ChainJob& j : jobs) flops += (lds ? 16.0 : 64.0) * (double)n * n * n * j.nsteps;
    {
        // per region: the bottom block row (16 n^3 flop) or the whole 2n x 2n product (64 n^3 flop)
        TimedLaunch tl(ctx, "chain", flops);
        if (lds) {
            const ChainSrc src{P, Pinv, disks ? disks->G : nullptr, disks ? disks->Ginv : nullptr,
                               disks ? disks->ph : nullptr, disks ? disks->phinv : nullptr};
            launch_chain_lds(s, nj, djobs.as<ChainJob>(), dsteps.as<ChainStep>(), src, n, e0, reflect, boost, refl);
        } else {
            if (disks) throw Error(BF_ERR_UNSUPPORTED, "disk steps need the LDS chain");
            hipLaunchKernelGGL(k_chain, dim3((unsigned)nj), dim3(256), 0, s, djobs.as<ChainJob>(),
                               dsteps.as<ChainStep>(), P, Pinv, n, e0, reflect, Tb.as<zd>(), cw.as<zd>(), boost,
                               refl);
        }
        BF_CHECK_LAUNCH();
    }
    if (!st) BF_HIP(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------------------
// Allocation-free chain (bf_transformer_chain_noalloc)
// ---------------------------------------------------------------------------
// The entry point's working memory (inverses, selection list, chain jobs and
// steps) is one fixed array per GPU, held from library load to process end.
// Host code never passes `scratch` itself (its name is a host-side stand-in):
// the array's device address on each GPU comes from hipGetSymbolAddress at
// context creation (noalloc_init, device current) and is kept per device.
static __device__ __align__(256) unsigned char scratch[BF_NOALLOC_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[BF_NOALLOC_MAX_DEVICES] = {};

void noalloc_init(int device) {
    if (device < 0 || device >= BF_NOALLOC_MAX_DEVICES) throw Error(BF_ERR_UNSUPPORTED, "device index above 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void* p = nullptr;
    BF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error(BF_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[device] = p;
}

static inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// bytes of scratch a call needs: inverses, selection list, jobs, steps
size_t noalloc_bytes(int nreg, int n, int nfreq) {
    const size_t slots = (size_t)nfreq * nreg, nsel = (size_t)nfreq * (nreg - 1);
    return align256(sizeof(zd) * (size_t)n * n * slots) + align256(sizeof(int) * nsel) +
           align256(sizeof(ChainJob) * (size_t)nfreq) + align256(sizeof(ChainStep) * (size_t)(nreg - 1));
}

// one chain per frequency over regions nreg-1 .. 1; sel lists the matrices to invert
__global__ void k_noalloc_jobs(ChainJob* jobs, int* sel, int nfreq, int nreg, int n) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nfreq) return;
    ChainJob j;
    j.step0 = 0;
    j.pbase = (int64_t)f * nreg;
    j.out = (int64_t)f * n;
    j.e0off = 0;
    j.nsteps = nreg - 1;
    j.pad = 0;
    jobs[f] = j;
    for (int r = 1; r < nreg; r++) sel[(size_t)f * (nreg - 1) + (r - 1)] = f * nreg + r;
}

void transformer_chain_noalloc(bf_ctx* ctx, int nreg, const double* eps_h, int n, int nfreq, const zd* P,
                               const zd* e0, const zd* reflect, zd* boost, zd* refl) {
    hipStream_t s = ctx->stream;
    unsigned char* base = ctx->device >= 0 && ctx->device < BF_NOALLOC_MAX_DEVICES
                              ? static_cast<unsigned char*>(g_scratch_ptr[ctx->device])
                              : nullptr;
    if (!base) throw Error(BF_ERR_HIP, "the scratch array of this device has no address (context not initialised)");
    if (n > INV_WAVE_MAX || n > CHAIN_LDS_MAX) throw Error(BF_ERR_UNSUPPORTED, "n above 48");
    if (noalloc_bytes(nreg, n, nfreq) > (size_t)BF_NOALLOC_SCRATCH_BYTES)
        throw Error(BF_ERR_UNSUPPORTED, "call exceeds the scratch array");
    const size_t nn = (size_t)n * n, slots = (size_t)nfreq * nreg, nsel = (size_t)nfreq * (nreg - 1);
    zd* Pinv = reinterpret_cast<zd*>(base);
    int* sel = reinterpret_cast<int*>(base + align256(sizeof(zd) * nn * slots));
    ChainJob* jobs = reinterpret_cast<ChainJob*>(reinterpret_cast<unsigned char*>(sel) + align256(sizeof(int) * nsel));
    ChainStep* steps =
        reinterpret_cast<ChainStep*>(reinterpret_cast<unsigned char*>(jobs) + align256(sizeof(ChainJob) * nfreq));
    // steps from the permittivities (bdry.eps[isnan] .= 1e30, src/boost_fractor_transformer.jl:289), staged
    // from this stack frame; the stream is synchronised before it returns
    zd eps[BF_NOALLOC_MAX_REGIONS];
    ChainStep hsteps[BF_NOALLOC_MAX_REGIONS - 1];
    for (int r = 0; r < nreg; r++) {
        eps[r] = cmk<double>(eps_h[2 * r], eps_h[2 * r + 1]);
        if (std::isnan(eps[r].re) || std::isnan(eps[r].im)) eps[r] = cmk<double>(1e30, 0.0);
    }
    for (int q = 0; q < nreg - 1; q++) hsteps[q] = chain_step(eps, nreg, q);
    BF_HIP(hipMemcpyAsync(steps, hsteps, sizeof(ChainStep) * (nreg - 1), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_noalloc_jobs, dim3((unsigned)((nfreq + 63) / 64)), dim3(64), 0, s, jobs, sel, nfreq, nreg,
                       n);
    BF_CHECK_LAUNCH();
    launch_inverse(ctx, P, Pinv, n, sel, nsel);
    launch_chain_lds(s, (size_t)nfreq, jobs, steps, ChainSrc{P, Pinv, nullptr, nullptr, nullptr, nullptr}, n, e0,
                     reflect, boost, refl);
    BF_HIP(hipStreamSynchronize(s));
}

void transformer_dev(bf_ctx* ctx, const bf_setup* st, const bf_coords* co, const bf_modes* modes, const double* freqs,
                     int nfreq, int kind, const double* prop_mats, double diskR, const double* emit,
                     const double* reflect, int precision, double* boost, double* refl) {
    (void)precision;
    hipStream_t s = ctx->stream;
    Grid g;
    DevBuf surf;
    std::vector<bool> flat;
    setup_common(ctx, st, co, g, surf, flat);
    ModesDev md;
    md.upload(modes, g.K(), s);
    const int n = md.n, nreg = st->nreg;
    const size_t nn = (size_t)n * n;
    std::vector<zd> eps = eps_fixed(st);
    DevBuf P(sizeof(zd) * nn * nreg * nfreq), Pinv(sizeof(zd) * nn * nreg * nfreq);
    if (prop_mats) {
        upload(P.p, prop_mats, sizeof(zd) * nn * nreg * nfreq, s);
    } else {
        std::vector<int> regs;
        for (int r = 1; r < nreg; r++) regs.push_back(r);  // regions 2..Nreg (:311-312)
        propagation_matrices_core(ctx, g, md, kind, diskR, nreg, st->distance, eps.data(), st->tilt_x, st->tilt_y,
                                  surf.as<zd>(), flat, freqs, nfreq, regs, P.as<zd>());
    }
    std::vector<int> sel;
    std::vector<ChainJob> jobs;
    std::vector<ChainStep> steps;
    append_chain_steps(eps, nreg, steps);  // the same region sequence at every frequency
    for (int f = 0; f < nfreq; f++) {
        for (int r = 1; r < nreg; r++) sel.push_back(f * nreg + r);
        ChainJob j;
        j.step0 = 0;
        j.nsteps = nreg - 1;
        j.pbase = (int64_t)f * nreg;
        j.out = (int64_t)f * n;
        j.e0off = 0;
        j.pad = 0;
        jobs.push_back(j);
    }
    DevBuf de0(sizeof(zd) * n), dref, dboost(sizeof(zd) * n * nfreq), drefl;
    upload(de0.p, emit, sizeof(zd) * n, s);
    if (reflect) {
        dref.alloc(sizeof(zd) * n);
        drefl.alloc(sizeof(zd) * n * nfreq);
        upload(dref.p, reflect, sizeof(zd) * n, s);
    }
    inverse_and_chains(ctx, n, P.as<zd>(), Pinv.as<zd>(), sel, jobs, steps, de0.as<zd>(), dref.as<zd>(),
                       dboost.as<zd>(), drefl.as<zd>());
    download(boost, dboost.p, sizeof(zd) * n * nfreq, s);
    if (reflect) download(refl, drefl.p, sizeof(zd) * n * nfreq, s);
    BF_HIP(hipStreamSynchronize(s));
}

}  // namespace bf

namespace bf {

bf_transformer_plan* transformer_plan_create(bf_ctx* ctx, const bf_setup* st, const bf_coords* co,
                                             const bf_modes* modes, const double* freqs, int nfreq, int kind,
                                             double diskR, const double* emit, const double* reflect) {
    hipStream_t s = ctx->stream;
    auto* pl = new bf_transformer_plan;
    try {
        pl->ctx = ctx;
        setup_common(ctx, st, co, pl->g, pl->surf, pl->flat);
        pl->md.upload(modes, pl->g.K(), s);
        pl->nreg = st->nreg;
        pl->nf = nfreq;
        pl->kind = kind;
        pl->diskR = diskR;
        pl->distance.assign(st->distance, st->distance + st->nreg);
        pl->tx.assign(st->tilt_x, st->tilt_x + st->nreg);
        pl->ty.assign(st->tilt_y, st->tilt_y + st->nreg);
        pl->freqs.assign(freqs, freqs + nfreq);
        pl->eps = eps_fixed(st);
        const int n = pl->md.n;
        const size_t nn = (size_t)n * n;
        pl->P.alloc(sizeof(zd) * nn * pl->nreg * nfreq);
        pl->Pinv.alloc(sizeof(zd) * nn * pl->nreg * nfreq);
        pl->de0.alloc(sizeof(zd) * n);
        upload(pl->de0.p, emit, sizeof(zd) * n, s);
        pl->dboost.alloc(sizeof(zd) * n * nfreq);
        pl->has_reflect = reflect != nullptr;
        if (reflect) {
            pl->dref.alloc(sizeof(zd) * n);
            pl->drefl.alloc(sizeof(zd) * n * nfreq);
            upload(pl->dref.p, reflect, sizeof(zd) * n, s);
        }
        BF_HIP(hipStreamSynchronize(s));
    } catch (...) {
        delete pl;
        throw;
    }
    return pl;
}

void transformer_plan_run(bf_transformer_plan* pl) {
    bf_ctx* ctx = pl->ctx;
    const int n = pl->md.n, nreg = pl->nreg, nf = pl->nf;
    pl->mc.reset_data();
    std::vector<int> regs;
    for (int r = 1; r < nreg; r++) regs.push_back(r);
    propagation_matrices_core(ctx, pl->g, pl->md, pl->kind, pl->diskR, nreg, pl->distance.data(), pl->eps.data(),
                              pl->tx.data(), pl->ty.data(), pl->surf.as<zd>(), pl->flat, pl->freqs.data(), nf, regs,
                              pl->P.as<zd>(), &pl->mc);
    std::vector<int> sel;
    std::vector<ChainJob> jobs;
    std::vector<ChainStep> steps;
    append_chain_steps(pl->eps, nreg, steps);  // the same region sequence at every frequency
    for (int f = 0; f < nf; f++) {
        for (int r = 1; r < nreg; r++) sel.push_back(f * nreg + r);
        jobs.push_back(ChainJob{0, (int64_t)f * nreg, (int64_t)f * n, 0, nreg - 1, 0});
    }
    inverse_and_chains(ctx, n, pl->P.as<zd>(), pl->Pinv.as<zd>(), sel, jobs, steps, pl->de0.as<zd>(),
                       pl->dref.as<zd>(), pl->dboost.as<zd>(), pl->drefl.as<zd>());
}

void transformer_plan_fetch(bf_transformer_plan* pl, double* boost, double* refl) {
    hipStream_t s = pl->ctx->stream;
    const size_t bytes = sizeof(zd) * pl->md.n * pl->nf;
    if (boost) download(boost, pl->dboost.p, bytes, s);
    if (refl && pl->has_reflect) download(refl, pl->drefl.p, bytes, s);
    BF_HIP(hipStreamSynchronize(s));
}

void transformer_plan_destroy(bf_transformer_plan* pl) { delete pl; }

// Many geometries (each its own setup and mode set) at the same frequencies.
// Geometries are taken mode set by mode set, in chunks whose propagation
// matrices fit a memory budget; every chunk runs its matrices' inverses and
// all its (geometry, frequency) chains in one launch each.  The plan keeps the
// grid, the mode sets' patterns, their axion sources and the frequencies on
// the device, and the k-bin grouping of the grid once built; every run
// recomputes the pattern FFTs, group Grams and flat-disk Gram matrix of each
// mode set it uses (once per set and run).
// With ncoords = nsets, mode set `set` lives on its own grid co[set] (grids of any size in one plan): the
// grid enters only the set's pattern transforms, Gram tables and Parseval scale, and from there on every
// region is an n x n matrix.
bf_transformer_batch_plan* transformer_batch_plan_create(bf_ctx* ctx, const bf_coords* co, int ncoords, int nsets,
                                                         const bf_modes* modesets, const double* const* emits,
                                                         const double* freqs, int nfreq, int kind, double diskR) {
    hipStream_t s = ctx->stream;
    auto* pl = new bf_transformer_batch_plan;
    try {
        pl->ctx = ctx;
        pl->grids.resize((size_t)ncoords);
        for (int k = 0; k < ncoords; k++) pl->grids[(size_t)k].upload(&co[k], s);
        pl->md.resize(nsets);
        pl->de0.resize(nsets);
        pl->mc.resize(nsets);
        for (int set = 0; set < nsets; set++) {
            pl->md[set].upload(&modesets[set], pl->grid(set).K(), s);
            pl->de0[set].alloc(sizeof(zd) * pl->md[set].n);
            upload(pl->de0[set].p, emits[set], sizeof(zd) * pl->md[set].n, s);
        }
        pl->freqs.assign(freqs, freqs + nfreq);
        pl->kind = kind;
        pl->diskR = diskR;
        BF_HIP(hipStreamSynchronize(s));
    } catch (...) {
        delete pl;
        throw;
    }
    return pl;
}

// ---------------------------------------------------------------------------
// Flat chunks of the batch plan.  When every geometry of a chunk has flat,
// untilted regions (no surface map, zero tilts), every region's matrix at
// every frequency follows from frequency-independent descriptors: an air gap's
// from the spacing series of its bin (coefficients per frequency), a disk's as
// diag(exp(-1im*kz*dz)) G.  The kernels expand the descriptors over the
// frequencies (grid y), so the host's work per chunk grows with the chunk's
// regions, not with regions x frequencies.  The matrices, inverses and chains
// are the ones the general path forms (the same kernels' arithmetic).
// ---------------------------------------------------------------------------
struct FlatRegion {
    int64_t base;  // matrix index at frequency 0; frequency f adds f * stride
    int stride;    // the geometry's region count
    int slot;      // permittivity index (k0^2 table row)
    double dz;
};
struct FlatSlice {  // series evaluation: entries [first, first + count) of one bin, key `key`, frequency f
    int key, first, count, f;
};

// INV: the inverses from the inverse's coefficients (key.e) into P = inv(P)'s array
template <bool INV>
__global__ void __launch_bounds__(128) k_series_eval_flat(const FlatSlice* __restrict__ sl,
                                                          const FlatRegion* __restrict__ ent,
                                                          const SeriesKey* __restrict__ keys, zd* P, int nn) {
    const int e = blockIdx.x * 128 + threadIdx.x;
    if (e >= nn) return;
    const FlatSlice S = sl[blockIdx.y];
    const SeriesKey k = keys[S.key];
    const zd* coef = INV ? k.e : k.d;
    zd d[SERIES_QMAX];
#pragma unroll
    for (int p = 0; p < SERIES_QMAX; p++) d[p] = p < k.q ? coef[(size_t)p * nn + e] : cmk<double>(0.0, 0.0);
    for (int i = 0; i < S.count; i++) {
        const FlatRegion R = ent[S.first + i];
        const double t = k.h > 0.0 ? (R.dz - k.c) / k.h : 0.0;
        zd acc = cmk<double>(0.0, 0.0);
#pragma unroll
        for (int p = SERIES_QMAX - 1; p >= 0; p--) acc = cmk<double>(d[p].re + t * acc.re, d[p].im + t * acc.im);
        P[(size_t)(R.base + (int64_t)S.f * R.stride) * nn + e] = acc;
    }
}

// flat disks, block (entry, frequency): P = diag(exp(-1im*kz*dz)) G and, with Ginv, inv(P) = Ginv diag(1/ph)
__global__ void __launch_bounds__(256) k_disk_flat(const FlatRegion* __restrict__ ent, const zd* __restrict__ k0sq,
                                                   int nf, const zd* __restrict__ G, const zd* __restrict__ Ginv,
                                                   const zd* __restrict__ kt, int n, zd* P, zd* Pinv) {
    __shared__ zd sph[BF_MAX_MODES];
    const FlatRegion R = ent[blockIdx.x];
    const int f = blockIdx.y;
    const zd k2 = k0sq[(size_t)R.slot * nf + f];
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const zd kz = csqrt_j(csub(k2, cmul(kt[i], kt[i])));
        sph[i] = cexp_j(rmul(R.dz, neg_i(kz)));
    }
    __syncthreads();
    const size_t off = (size_t)(R.base + (int64_t)f * R.stride) * n * n;
    each2d(threadIdx.x, blockDim.x, n, n, [&](int j, int i) {
        const int e = i + n * j;
        P[off + e] = cmul(sph[i], G[e]);
        if (Ginv) {
            const zd a = Ginv[e], d = sph[j];
            const double den = d.re * d.re + d.im * d.im;
            Pinv[off + e] = cmk<double>((a.re * d.re + a.im * d.im) / den, (a.im * d.re - a.re * d.im) / den);
        }
    });
}

// flat disks formed inside the chain (ChainSrc): thread (entry x, frequency f, mode i) writes the phase
// exp(-1im*kz_i*dz) of the entry's matrix slot m = base + f stride at ph[m n + i], as k_disk_flat forms it,
// and its reciprocal at phinv[m n + i]
__global__ void __launch_bounds__(256) k_disk_phases(const FlatRegion* __restrict__ ent, size_t total,
                                                     const zd* __restrict__ k0sq, int nf, const zd* __restrict__ kt,
                                                     int n, zd* ph, zd* phinv) {
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= total) return;
    const int i = (int)(t % n);
    const size_t xf = t / n;
    const int f = (int)(xf % nf);
    const FlatRegion R = ent[xf / nf];
    const zd k2 = k0sq[(size_t)R.slot * nf + f];
    const zd kz = csqrt_j(csub(k2, cmul(kt[i], kt[i])));
    const zd d = cexp_j(rmul(R.dz, neg_i(kz)));
    const double den = d.re * d.re + d.im * d.im;
    const size_t o = (size_t)(R.base + (int64_t)f * R.stride) * n + i;
    ph[o] = d;
    phinv[o] = cmk<double>(d.re / den, -d.im / den);
}

// true if the flat chunk path serves geometries confs[c0, c1)
static bool flat_chunk_ok(const bf_transformer_batch_plan* pl, const ModesDev& md, const bf_setup* setups,
                          const std::vector<int>& confs, size_t c0, size_t c1) {
    if (pl->kind != BF_PROP_PROPAGATOR && pl->kind != BF_PROP_NOTILTS) return false;
    if (!opts().parseval_groups || !opts().parseval_series) return false;
    if (md.n > INV_WAVE_MAX || md.n > CHAIN_LDS_MAX) return false;
    for (size_t q = c0; q < c1; q++) {
        const bf_setup* st = &setups[confs[q]];
        if (st->surfaces) return false;
        for (int r = 0; r < st->nreg; r++)
            if (st->tilt_x[r] != 0.0 || st->tilt_y[r] != 0.0 || !std::isfinite(st->distance[r])) return false;
    }
    return true;
}

// The host side of a flat chunk (flat_prepare) and its launches (flat_launch).
struct FlatWork {
    std::vector<FlatRegion> ent, disks, inv;  // gaps by group; disks; regions whose inverses take the LU
    std::vector<zd> dk0;                      // disks' k0^2 [permittivity][frequency]
    std::vector<SeriesKey> fresh, keys;
    std::vector<SeriesKeyId> fresh_key, key_ids;
    std::vector<FlatSlice> slices, islices;  // series evaluations of the matrices and of the inverses
    std::vector<ChainStep> steps;
    std::vector<ChainJob> jobs;
    ParsevalQueue direct;  // gaps whose (frequency, bin) the series does not serve
    int nrows = 0;
    bool ginv = false;
    int* bad = nullptr;  // device flag of the inverse series (null: every gap's inverse by the LU)
    int64_t nmats = 0;   // matrix slots of the chunk
    std::vector<char> step_disk;  // per chain step: its region is a disk
};

// Host preparation of geometries confs[c0, c1) (matrices at P, chunk-relative indices).  False if the
// mode set's tables do not allow the flat path (then nothing was queued).
static bool flat_prepare(bf_ctx* ctx, bf_transformer_batch_plan* pl, int set, ModeCache& mc, const bf_setup* setups,
                         const std::vector<int>& confs, size_t c0, size_t c1, const int64_t* boost_off, zd* P,
                         FlatWork& fw) {
    const ModesDev& md = pl->md[set];
    const int n = md.n, nf = (int)pl->freqs.size();
    const size_t nn = (size_t)n * n;
    // regions: air gaps (Parseval) and disks, permittivities by bits
    std::vector<FlatRegion> gaps;
    std::vector<zd> gap_eps, disk_eps;
    auto eps_index = [](std::vector<zd>& list, const zd& e) {
        for (size_t i = 0; i < list.size(); i++)
            if (std::memcmp(&list[i], &e, sizeof(zd)) == 0) return (int)i;
        list.push_back(e);
        return (int)list.size() - 1;
    };
    int64_t base = 0;
    for (size_t q = c0; q < c1; q++) {
        const bf_setup* sq = &setups[confs[q]];
        const int nreg = sq->nreg;
        const std::vector<zd> eps = eps_fixed(sq);
        for (int r = 1; r < nreg; r++) {
            zd e = eps[r];
            if (e.re == 1.0) {
                if (pl->kind == BF_PROP_PROPAGATOR && e.im == 0.0) e.im = -0.0;
                gaps.push_back(FlatRegion{base + r, nreg, eps_index(gap_eps, e), sq->distance[r]});
            } else {
                fw.disks.push_back(FlatRegion{base + r, nreg, eps_index(disk_eps, e), sq->distance[r]});
            }
        }
        const int64_t step0 = (int64_t)fw.steps.size();
        append_chain_steps(eps, nreg, fw.steps);
        for (size_t k = (size_t)step0; k < fw.steps.size(); k++)
            fw.step_disk.push_back(eps[fw.steps[k].reg].re != 1.0);
        for (int f = 0; f < nf; f++) {
            ChainJob j;
            j.step0 = step0;
            j.nsteps = nreg - 1;
            j.pbase = base + (int64_t)f * nreg;
            j.out = boost_off[confs[q]] + (int64_t)f * n;  // the plan's results, geometry-major
            j.e0off = 0;
            j.pad = 0;
            fw.jobs.push_back(j);
        }
        base += (int64_t)nreg * nf;
    }
    const bool grouped = mode_tables(ctx, pl->grid(set), md, mc, !fw.disks.empty(), !gaps.empty(), pl->diskR);
    if (!grouped && !gaps.empty()) return false;
    // k0^2 per permittivity and frequency
    auto k0_table = [&](const std::vector<zd>& list) {
        std::vector<zd> t(list.size() * (size_t)nf);
        for (size_t v = 0; v < list.size(); v++)
            for (int f = 0; f < nf; f++) {
                const zd k0 = k0_of(list[v], 299792458.0 / pl->freqs[(size_t)f]);
                t[v * nf + f] = cmul(k0, k0);
            }
        return t;
    };
    const std::vector<zd> gk0 = k0_table(gap_eps);
    fw.dk0 = k0_table(disk_eps);
    fw.nmats = base;
    fw.ginv = !fw.disks.empty() && ginv_usable(ctx, mc, md);
    if (!fw.ginv) fw.inv = fw.disks;
    // with inv(G), the chain forms each disk's matrix and inverse from G, inv(G) and its phases
    if (fw.ginv)
        for (size_t k = 0; k < fw.steps.size(); k++) fw.steps[k].disk = fw.step_disk[k];
    // the gaps by (permittivity, bin), dz = 0 apart; their keys per frequency: formed earlier in the run,
    // formed now, or (too few gaps) the direct product
    if (mc.series_w <= 0.0 && !gk0.empty()) {
        double k2 = 0.0;
        for (const zd& v : gk0) k2 = std::max(k2, std::hypot(v.re, v.im));
        mc.series_w = 2.0 / std::sqrt(4.0 * k2 + mc.gk2max);
    }
    const double w = mc.series_w;
    std::map<std::pair<int, int64_t>, std::vector<int>> groups;  // (permittivity, bin; zero: INT64_MIN) -> gaps
    for (size_t i = 0; i < gaps.size(); i++) {
        const double dz = gaps[i].dz;
        const int64_t b =
            dz == 0.0 ? INT64_MIN : (dz > 0.0 && dz / w < 1e15 ? (int64_t)std::floor(dz / w) : INT64_MAX);
        groups[{gaps[i].slot, b}].push_back((int)i);
    }
    auto bits = [](double v) {
        uint64_t b;
        std::memcpy(&b, &v, 8);
        return b;
    };
    std::map<SeriesKeyId, int> key_index;
    std::vector<int> kf((size_t)nf);
    for (const auto& gkv : groups) {
        const int v = gkv.first.first;
        const int64_t b = gkv.first.second;
        const int first = (int)fw.ent.size(), cnt = (int)gkv.second.size();
        for (int i : gkv.second) fw.ent.push_back(gaps[(size_t)i]);
        bool inv_all = fw.bad != nullptr;  // every frequency's inverses from the series
        for (int f = 0; f < nf; f++) {
            const zd k0 = gk0[(size_t)v * nf + f];
            int ki = -1;
            if (b != INT64_MAX && k0.im <= 0.0 && std::isfinite(k0.re)) {
                const SeriesKeyId id = b == INT64_MIN ? SeriesKeyId{0, 0, SERIES_ZERO_BIN}
                                                      : SeriesKeyId{bits(k0.re), bits(k0.im), (uint64_t)b};
                auto it = key_index.find(id);
                if (it != key_index.end()) {
                    ki = it->second;
                } else {
                    auto c = mc.series.find(id);
                    SeriesKey sk;
                    bool ok = false;
                    if (c != mc.series.end()) {
                        sk = SeriesKey{k0, c->second.c, c->second.h, c->second.q, 0, c->second.d, c->second.e};
                        ok = true;
                    } else if (series_key_for(mc, id, k0, b == INT64_MIN ? cnt * nf : cnt, sk)) {
                        sk.row0 = fw.nrows;
                        fw.nrows += sk.q;
                        fw.fresh.push_back(sk);
                        fw.fresh_key.push_back(id);
                        ok = true;
                    }
                    if (ok) {
                        ki = (int)fw.keys.size();
                        key_index[id] = ki;
                        fw.keys.push_back(sk);
                        fw.key_ids.push_back(id);
                    }
                }
            }
            kf[(size_t)f] = ki;
            // keys formed in this chunk get their inverse coefficients; earlier ones have them if they
            // were formed by a flat chunk of this run
            if (ki < 0 || (fw.keys[(size_t)ki].d && !fw.keys[(size_t)ki].e)) inv_all = false;
            if (ki >= 0) {
                for (int o = 0; o < cnt; o += SERIES_SLICE)
                    fw.slices.push_back(FlatSlice{ki, first + o, std::min(SERIES_SLICE, cnt - o), f});
            } else {
                for (int i = first; i < first + cnt; i++) {
                    const FlatRegion& R = fw.ent[(size_t)i];
                    fw.direct.pd.push_back(PwDesc{k0, R.dz, 0, 0});
                    fw.direct.dst.push_back(P + (size_t)(R.base + (int64_t)f * R.stride) * nn);
                }
            }
        }
        if (inv_all) {
            for (int f = 0; f < nf; f++)
                for (int o = 0; o < cnt; o += SERIES_SLICE)
                    fw.islices.push_back(FlatSlice{kf[(size_t)f], first + o, std::min(SERIES_SLICE, cnt - o), f});
        } else {
            fw.inv.insert(fw.inv.end(), fw.ent.begin() + first, fw.ent.end());
        }
    }
    return true;
}

// The launches of a prepared flat chunk (uploads through the plan's stager)
static void flat_launch(bf_ctx* ctx, bf_transformer_batch_plan* pl, int set, ModeCache& mc, FlatWork& fw, zd* P,
                        zd* Pinv, zd* dboost) {
    hipStream_t s = ctx->stream;
    const ModesDev& md = pl->md[set];
    const int n = md.n, nf = (int)pl->freqs.size();
    const size_t nn = (size_t)n * n;
    const double scale = 1.0 / (double)pl->grid(set).K();
    Stager* st = &pl->stager;
    // series coefficients of the new keys, then every key's device pointer
    series_form(ctx, mc, n, scale, fw.fresh, fw.fresh_key, fw.nrows, st, fw.bad);
    for (size_t k = 0; k < fw.keys.size(); k++)
        if (!fw.keys[k].d) {
            const ModeCache::SeriesEntry& E = mc.series[fw.key_ids[k]];
            fw.keys[k].d = E.d;
            fw.keys[k].e = E.e;
        }
    if (!fw.slices.empty()) {
        DevBuf& dent = scratch_buf(ctx, SB_FENT, sizeof(FlatRegion) * fw.ent.size());
        DevBuf& dkey = scratch_buf(ctx, SB_SD, sizeof(SeriesKey) * fw.keys.size());
        DevBuf& dsl = scratch_buf(ctx, SB_SSLICE, sizeof(FlatSlice) * fw.slices.size());
        stage_upload(st, dent.p, fw.ent.data(), sizeof(FlatRegion) * fw.ent.size(), s);
        stage_upload(st, dkey.p, fw.keys.data(), sizeof(SeriesKey) * fw.keys.size(), s);
        stage_upload(st, dsl.p, fw.slices.data(), sizeof(FlatSlice) * fw.slices.size(), s);
        TimedLaunch tl(ctx, "pseries", 0.0);
        for (size_t s0 = 0; s0 < fw.slices.size(); s0 += 65535) {
            const size_t ns = std::min<size_t>(65535, fw.slices.size() - s0);
            hipLaunchKernelGGL(k_series_eval_flat<false>, dim3((unsigned)((nn + 127) / 128), (unsigned)ns), dim3(128),
                               0, s, dsl.as<FlatSlice>() + s0, dent.as<FlatRegion>(), dkey.as<SeriesKey>(), P,
                               (int)nn);
            BF_CHECK_LAUNCH();
        }
        if (!fw.islices.empty()) {
            DevBuf& disl = scratch_buf(ctx, SB_FISLICE, sizeof(FlatSlice) * fw.islices.size());
            stage_upload(st, disl.p, fw.islices.data(), sizeof(FlatSlice) * fw.islices.size(), s);
            for (size_t s0 = 0; s0 < fw.islices.size(); s0 += 65535) {
                const size_t ns = std::min<size_t>(65535, fw.islices.size() - s0);
                hipLaunchKernelGGL(k_series_eval_flat<true>, dim3((unsigned)((nn + 127) / 128), (unsigned)ns),
                                   dim3(128), 0, s, disl.as<FlatSlice>() + s0, dent.as<FlatRegion>(),
                                   dkey.as<SeriesKey>(), Pinv, (int)nn);
                BF_CHECK_LAUNCH();
            }
        }
    }
    parseval_flush(ctx, mc, n, scale, fw.direct, st);
    ChainSrc disk_src{P, Pinv, mc.G.as<zd>(), fw.ginv ? mc.Ginv.as<zd>() : nullptr, nullptr, nullptr};
    if (!fw.disks.empty()) {
        DevBuf& dd = scratch_buf(ctx, SB_FDISK, sizeof(FlatRegion) * fw.disks.size());
        DevBuf& dk = scratch_buf(ctx, SB_FK0, sizeof(zd) * fw.dk0.size());
        stage_upload(st, dd.p, fw.disks.data(), sizeof(FlatRegion) * fw.disks.size(), s);
        stage_upload(st, dk.p, fw.dk0.data(), sizeof(zd) * fw.dk0.size(), s);
        if (fw.ginv) {
            // phases only: the chain forms the disks' matrices and inverses from G and inv(G)
            DevBuf& dph = scratch_buf(ctx, SB_FPH, 2 * sizeof(zd) * (size_t)fw.nmats * n);
            zd* ph = dph.as<zd>();
            zd* phinv = ph + (size_t)fw.nmats * n;
            disk_src.ph = ph;
            disk_src.phinv = phinv;
            const size_t total = fw.disks.size() * (size_t)nf * n;
            hipLaunchKernelGGL(k_disk_phases, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                               dd.as<FlatRegion>(), total, dk.as<zd>(), nf, md.kt.as<zd>(), n, ph, phinv);
            BF_CHECK_LAUNCH();
        } else {
            for (size_t d0 = 0; d0 < fw.disks.size(); d0 += 65535) {
                const size_t nd = std::min<size_t>(65535, fw.disks.size() - d0);
                hipLaunchKernelGGL(k_disk_flat, dim3((unsigned)nd, (unsigned)nf), dim3(256), 0, s,
                                   dd.as<FlatRegion>() + d0, dk.as<zd>(), nf, mc.G.as<zd>(), nullptr,
                                   md.kt.as<zd>(), n, P, Pinv);
                BF_CHECK_LAUNCH();
            }
        }
    }
    if (!fw.inv.empty()) {
        std::vector<InvSel> bs(fw.inv.size());
        for (size_t i = 0; i < fw.inv.size(); i++) bs[i] = InvSel{fw.inv[i].base, fw.inv[i].stride};
        DevBuf& di = scratch_buf(ctx, SB_FINV, sizeof(InvSel) * bs.size());
        stage_upload(st, di.p, bs.data(), sizeof(InvSel) * bs.size(), s);
        TimedLaunch tl(ctx, "inverse", (8.0 / 3.0 + 8.0) * (double)n * n * n * (double)bs.size() * nf);
        const size_t smem = sizeof(zd) * nn;
        allow_lds((const void*)k_batched_inverse_wave, smem);
        for (size_t i0 = 0; i0 < bs.size(); i0 += 65535) {
            const size_t ni = std::min<size_t>(65535, bs.size() - i0);
            hipLaunchKernelGGL(k_batched_inverse_wave, dim3((unsigned)ni, (unsigned)nf), dim3(64), smem, s, P, Pinv,
                               n, (const int*)nullptr, di.as<InvSel>() + i0);
            BF_CHECK_LAUNCH();
        }
    }
    inverse_and_chains(ctx, n, P, Pinv, std::vector<int>(), fw.jobs, fw.steps, pl->de0[set].as<zd>(), nullptr, dboost,
                       nullptr, st, disk_src.ph ? &disk_src : nullptr);
}

// one run; with `bad`, flat chunks take the gaps' inverses from the inverse series and raise *bad if one
// of them did not converge
static void plan_run_once(bf_transformer_batch_plan* pl, int nconf, const bf_setup* setups, const int32_t* set_of,
                          const int64_t* boost_off, int* bad) {
    bf_ctx* ctx = pl->ctx;
    hipStream_t s = ctx->stream;
    const int nfreq = (int)pl->freqs.size();
    const double* freqs = pl->freqs.data();
    // the spacing-series coefficients depend on the geometries' spacings: formed anew in every run
    BF_HIP(hipStreamSynchronize(s));
    for (ModeCache& c : pl->mc) c.reset_series();
    // results: geometry c's [nfreq][n] block at complex element boost_off[c]
    size_t total = 0;
    pl->blocks.clear();
    for (int c = 0; c < nconf; c++) {
        const size_t len = (size_t)pl->md[set_of[c]].n * nfreq;
        pl->blocks.push_back({(size_t)boost_off[c], len});
        total = std::max(total, (size_t)boost_off[c] + len);
    }
    if (pl->dboost.bytes < sizeof(zd) * total) pl->dboost.alloc(sizeof(zd) * total);
    pl->nboost = total;
    // bytes of P + Pinv per chunk (option batch_bytes overrides, e.g. to exercise chunking on small cases)
    const size_t budget = opts().batch_bytes > 0 ? (size_t)opts().batch_bytes : BF_DEFAULT_BATCH_BYTES;
    bool first_chunk = true;
    // with timing on: the host's own work per chunk (lists, descriptors, launches) and its waits for a
    // staging slot, as pseudo kernel classes "host_prep" and "host_wait"
    using clk = std::chrono::steady_clock;
    auto ms_since = [](clk::time_point a, clk::time_point b) {
        return std::chrono::duration<double, std::milli>(b - a).count();
    };
    double host_ms = 0.0, wait_ms = 0.0, geo_ms = 0.0;
    int64_t nchunks = 0;
    for (int set = 0; set < (int)pl->md.size(); set++) {
        std::vector<int> confs;
        for (int c = 0; c < nconf; c++)
            if (set_of[c] == set) confs.push_back(c);
        if (confs.empty()) continue;
        const ModesDev& md = pl->md[set];
        const Grid& g = pl->grid(set);
        // the mode set's pattern transforms, k-bin Gram tables and flat-disk Gram depend only on the plan's
        // modes and grid: formed by the first run that needs them and kept (no wait between mode sets)
        ModeCache& cache = pl->mc[set];
        const int n = md.n;
        const size_t nn = (size_t)n * n;
        for (size_t c0 = 0; c0 < confs.size();) {
            const auto h0 = clk::now();
            // chunk [c0, c1): sum of nreg * nfreq matrices within the budget (at least one geometry); the
            // run's first chunk holds a few geometries only (at most 256 MiB of matrices), so the device
            // starts while the host prepares more
            const size_t chunk_budget = first_chunk ? std::min(budget / 16, (size_t)256 << 20) : budget;
            size_t c1 = c0, mats = 0;
            while (c1 < confs.size()) {
                const size_t m = (size_t)setups[confs[c1]].nreg * nfreq;
                if (c1 > c0 && (mats + m) * nn * sizeof(zd) * 2 > chunk_budget) break;
                mats += m;
                c1++;
            }
            first_chunk = false;
            DevBuf& P = scratch_buf(ctx, SB_P, sizeof(zd) * nn * mats);
            DevBuf& Pinv = scratch_buf(ctx, SB_PINV, sizeof(zd) * nn * mats);
            // chunks of flat, untilted geometries: frequency-independent descriptors expanded on the device
            FlatWork fw;
            fw.bad = bad;
            if (flat_chunk_ok(pl, md, setups, confs, c0, c1) &&
                flat_prepare(ctx, pl, set, cache, setups, confs, c0, c1, boost_off, P.as<zd>(), fw)) {
                const auto h1 = clk::now();
                pl->stager.begin_chunk();
                const auto h2 = clk::now();
                flat_launch(ctx, pl, set, cache, fw, P.as<zd>(), Pinv.as<zd>(), pl->dboost.as<zd>());
                pl->stager.end_chunk(s);
                const auto h3 = clk::now();
                host_ms += ms_since(h0, h1) + ms_since(h2, h3);
                geo_ms += ms_since(h0, h1);
                wait_ms += ms_since(h1, h2);
                nchunks++;
                c0 = c1;
                continue;
            }
            std::vector<int> sel;
            std::vector<ChainJob> jobs;
            std::vector<ChainStep> steps;
            size_t base = 0;
            ParsevalQueue pq;
            pq.diag_base = P.as<zd>();
            for (size_t q = c0; q < c1; q++) {
                const bf_setup* st = &setups[confs[q]];
                const int nreg = st->nreg;
                DevBuf surf;
                std::vector<bool> flat;
                upload_surfaces(ctx, st, g.K(), surf, flat);
                std::vector<zd> eps = eps_fixed(st);
                std::vector<int> regs;
                for (int r = 1; r < nreg; r++) regs.push_back(r);
                propagation_matrices_core(ctx, g, md, pl->kind, pl->diskR, nreg, st->distance, eps.data(),
                                          st->tilt_x, st->tilt_y, surf.as<zd>(), flat, freqs, nfreq, regs,
                                          P.as<zd>() + base * nn, &cache, &pq);
                // the geometry's chain steps do not depend on the frequency: one list for all its jobs
                const int64_t step0 = (int64_t)steps.size();
                append_chain_steps(eps, nreg, steps);
                for (int f = 0; f < nfreq; f++) {
                    for (int r = 1; r < nreg; r++) sel.push_back((int)(base + (size_t)f * nreg + r));
                    ChainJob j;
                    j.step0 = step0;
                    j.nsteps = nreg - 1;
                    j.pbase = (int64_t)(base + (size_t)f * nreg);
                    j.out = boost_off[confs[q]] + (int64_t)f * n;  // the plan's results, geometry-major
                    j.e0off = 0;
                    j.pad = 0;
                    jobs.push_back(j);
                }
                base += (size_t)nreg * nfreq;
            }
            // the chunk's uploads go through pinned staging, its kernels queue behind the previous chunk's
            // (every device buffer is used in stream order), and the host goes on to prepare the next chunk
            const auto h1 = clk::now();
            pl->stager.begin_chunk();
            const auto h2 = clk::now();
            parseval_flush(ctx, cache, n, 1.0 / (double)g.K(), pq, &pl->stager);
            // flat disks taken from the Gram matrix get their inverses from inv(G) with the disk kernel
            const bool use_ginv = ginv_usable(ctx, cache, md);
            if (use_ginv) {
                std::vector<char> from_g(mats, 0);
                for (size_t i = 0; i < pq.dslot.size(); i++)
                    if (pq.dfrom[i]) from_g[(size_t)pq.dslot[i]] = 1;
                size_t keep = 0;
                for (size_t i = 0; i < sel.size(); i++)
                    if (!from_g[(size_t)sel[i]]) sel[keep++] = sel[i];
                sel.resize(keep);
            }
            disk_diag_flush(ctx, cache, md, pq, &pl->stager, use_ginv ? Pinv.as<zd>() : nullptr);
            inverse_and_chains(ctx, n, P.as<zd>(), Pinv.as<zd>(), sel, jobs, steps, pl->de0[set].as<zd>(), nullptr,
                               pl->dboost.as<zd>(), nullptr, &pl->stager);
            pl->stager.end_chunk(s);
            const auto h3 = clk::now();
            host_ms += ms_since(h0, h1) + ms_since(h2, h3);
            geo_ms += ms_since(h0, h1);
            wait_ms += ms_since(h1, h2);
            nchunks++;
            c0 = c1;
        }
    }
    if (ctx->timing) {
        bf_kernel_timer& hp = ctx->timers["host_prep"];
        hp.ms += host_ms;
        hp.launches += nchunks;
        bf_kernel_timer& hw = ctx->timers["host_wait"];
        hw.ms += wait_ms;
        hw.launches += nchunks;
        bf_kernel_timer& hg = ctx->timers["host_geometries"];  // the part of host_prep spent per geometry
        hg.ms += geo_ms;
        hg.launches += nchunks;
    }
    BF_HIP(hipStreamSynchronize(s));
}

// A run with the inverse series (option series_inverse) checks their convergence once at its end and, if
// a key's series fell short, is redone with every gap inverse by the LU.  series_inverse = 2 forces that
// path (tests).
void transformer_batch_plan_run(bf_transformer_batch_plan* pl, int nconf, const bf_setup* setups,
                                const int32_t* set_of, const int64_t* boost_off) {
    const int mode = opts().series_inverse;
    if (mode == 0) return plan_run_once(pl, nconf, setups, set_of, boost_off, nullptr);
    hipStream_t s = pl->ctx->stream;
    if (!pl->dbad.p) pl->dbad.alloc(sizeof(int));
    const int zero = 0;
    BF_HIP(hipMemcpyAsync(pl->dbad.p, &zero, sizeof(int), hipMemcpyHostToDevice, s));
    plan_run_once(pl, nconf, setups, set_of, boost_off, pl->dbad.as<int>());
    int bad = 0;
    BF_HIP(hipMemcpyAsync(&bad, pl->dbad.p, sizeof(int), hipMemcpyDeviceToHost, s));
    BF_HIP(hipStreamSynchronize(s));
    if (bad || mode == 2) {
        pl->reruns++;
        plan_run_once(pl, nconf, setups, set_of, boost_off, nullptr);
    }
}

// copies the last run's geometry blocks to the same offsets of `boost` (the
// elements between blocks are left as they are); runs of adjacent blocks go
// in one copy
void transformer_batch_plan_fetch(bf_transformer_batch_plan* pl, double* boost) {
    hipStream_t s = pl->ctx->stream;
    std::vector<std::pair<size_t, size_t>> b = pl->blocks;
    std::sort(b.begin(), b.end());
    for (size_t i = 0; i < b.size();) {
        size_t lo = b[i].first, hi = b[i].first + b[i].second;
        size_t j = i + 1;
        while (j < b.size() && b[j].first <= hi) {
            hi = std::max(hi, b[j].first + b[j].second);
            j++;
        }
        if (hi > lo) download(boost + 2 * lo, pl->dboost.as<zd>() + lo, sizeof(zd) * (hi - lo), s);
        i = j;
    }
    BF_HIP(hipStreamSynchronize(s));
}

void transformer_batch_plan_destroy(bf_transformer_batch_plan* pl) { delete pl; }

void transformer_batch_dev(bf_ctx* ctx, int nconf, const bf_setup* setups, const bf_coords* co, int nsets,
                           const bf_modes* modesets, const double* const* emits, const int32_t* set_of,
                           const double* freqs, int nfreq, int kind, double diskR, const int64_t* boost_off,
                           double* boost) {
    std::unique_ptr<bf_transformer_batch_plan> pl(
        transformer_batch_plan_create(ctx, co, 1, nsets, modesets, emits, freqs, nfreq, kind, diskR));
    transformer_batch_plan_run(pl.get(), nconf, setups, set_of, boost_off);
    transformer_batch_plan_fetch(pl.get(), boost);
}

// transformer_trace_back (:372-421): one block walks the regions applying G_s P_s
__global__ void k_trace_back(const ChainStep* __restrict__ steps, int nsteps, const zd* __restrict__ P,
                             const zd* __restrict__ Pinv, int n, const zd* __restrict__ sol0, int includes_axion,
                             const zd* __restrict__ emit, zd* sol, zd* out) {
    const int N2 = 2 * n;
    const size_t nn = (size_t)n * n;
    zd* cur = sol;
    zd* nxt = sol + N2;
    for (int i = threadIdx.x; i < N2; i += blockDim.x) cur[i] = sol0[i];
    __syncthreads();
    for (int q = 0; q < nsteps; q++) {
        const ChainStep st = steps[q];
        const zd* Pm = P + (size_t)st.reg * nn;
        const zd* Pi = Pinv + (size_t)st.reg * nn;
        // [a P, b Pi; b P, a Pi] * [x1; x2]
        for (int i = threadIdx.x; i < N2; i += blockDim.x) {
            const bool top = i < n;
            const int ii = top ? i : i - n;
            const zd ca = top ? st.a : st.b, cb = top ? st.b : st.a;
            zd s1 = cmk<double>(0.0, 0.0), s2 = cmk<double>(0.0, 0.0);
            for (int k = 0; k < n; k++) {
                s1 = cadd(s1, cmul(cmul(ca, Pm[ii + (size_t)n * k]), cur[k]));
                s2 = cadd(s2, cmul(cmul(cb, Pi[ii + (size_t)n * k]), cur[n + k]));
            }
            zd v = cadd(s1, s2);
            if (includes_axion) v = cadd(v, cmul(st.S, emit[ii]));
            nxt[i] = v;
        }
        __syncthreads();
        for (int i = threadIdx.x; i < N2; i += blockDim.x) {
            cur[i] = nxt[i];
            out[(size_t)q * N2 + i] = nxt[i];
        }
        __syncthreads();
    }
}

void trace_back_dev(bf_ctx* ctx, const double* reflected, const double* input, const bf_setup* st,
                    const bf_coords* co, const bf_modes* modes, double f, int kind, const double* prop_mats,
                    double diskR, int includes_axion, const double* emit, double* fields) {
    hipStream_t s = ctx->stream;
    Grid g;
    DevBuf surf;
    std::vector<bool> flat;
    setup_common(ctx, st, co, g, surf, flat);
    ModesDev md;
    md.upload(modes, g.K(), s);
    const int n = md.n, nreg = st->nreg, N2 = 2 * n;
    const size_t nn = (size_t)n * n;
    const zd* eps = reinterpret_cast<const zd*>(st->eps);  // used as given (no NaN fix in trace_back)
    DevBuf P(sizeof(zd) * nn * nreg), Pinv(sizeof(zd) * nn * nreg);
    if (prop_mats) {
        upload(P.p, prop_mats, sizeof(zd) * nn * nreg, s);
    } else {
        std::vector<int> regs;
        for (int r = 1; r < nreg; r++) regs.push_back(r);
        propagation_matrices_core(ctx, g, md, kind, diskR, nreg, st->distance, eps, st->tilt_x, st->tilt_y,
                                  surf.as<zd>(), flat, &f, 1, regs, P.as<zd>());
    }
    std::vector<int> sel;
    for (int r = 1; r < nreg; r++) sel.push_back(r);
    DevBuf dsel(sizeof(int) * sel.size());
    upload(dsel.p, sel.data(), sizeof(int) * sel.size(), s);
    launch_inverse(ctx, P.as<zd>(), Pinv.as<zd>(), n, dsel.as<int>(), sel.size());
    // steps s = 1 .. nreg-1: region idx_reg(s) = nreg-s+1 (1-based)
    const int nsteps = nreg - 1;
    std::vector<ChainStep> steps(nsteps);
    const zd one = cmk<double>(1.0, 0.0);
    for (int q = 0; q < nsteps; q++) {
        const int sidx = q + 1;
        const int ir = nreg - sidx, ir1 = nreg - sidx - 1;
        const zd nl = csqrt_j(eps[ir]), nr = csqrt_j(eps[ir1]);
        const zd sc = cdiv_h(one, rmul(2.0, nr));
        ChainStep c;
        c.S = rmul(0.5, csub(cdiv_h(one, cmul(nr, nr)), cdiv_h(one, cmul(nl, nl))));
        c.a = cmul(sc, cadd(nr, nl));
        c.b = cmul(sc, csub(nr, nl));
        c.reg = ir;
        c.disk = 0;
        steps[q] = c;
    }
    std::vector<zd> sol0(N2);
    const zd* in = reinterpret_cast<const zd*>(input);
    const zd* rf = reinterpret_cast<const zd*>(reflected);
    for (int i = 0; i < n; i++) {
        sol0[i] = in[i];
        sol0[n + i] = rf[i];
    }
    DevBuf dsteps(sizeof(ChainStep) * nsteps), dsol0(sizeof(zd) * N2), dsol(sizeof(zd) * 2 * N2),
        dout(sizeof(zd) * N2 * std::max(nsteps, 1)), dem(sizeof(zd) * n);
    upload(dsteps.p, steps.data(), sizeof(ChainStep) * nsteps, s);
    upload(dsol0.p, sol0.data(), sizeof(zd) * N2, s);
    if (emit) upload(dem.p, emit, sizeof(zd) * n, s);
    hipLaunchKernelGGL(k_trace_back, dim3(1), dim3(256), 0, s, dsteps.as<ChainStep>(), nsteps, P.as<zd>(),
                       Pinv.as<zd>(), n, dsol0.as<zd>(), includes_axion, dem.as<zd>(), dsol.as<zd>(), dout.as<zd>());
    BF_CHECK_LAUNCH();
    std::vector<zd> outs((size_t)N2 * std::max(nsteps, 1));
    download(outs.data(), dout.p, sizeof(zd) * N2 * nsteps, s);
    BF_HIP(hipStreamSynchronize(s));
    // fields_regions[idx_reg(1)] = initial solution, fields_regions[idx_reg(s+1)] = after step s
    // fields_regions[r, :, :] = solution (a 2n vector into the 2 x n slice, column-major):
    // element (r, d, m) = solution[d + 2m]; layout here is [nreg][2][n]
    zd* fo = reinterpret_cast<zd*>(fields);
    auto put = [&](int r, const zd* sol) {
        for (int d = 0; d < 2; d++)
            for (int m = 0; m < n; m++) fo[((size_t)r * 2 + d) * n + m] = sol[d + 2 * m];
    };
    put(nreg - 1, sol0.data());
    for (int q = 0; q < nsteps; q++) put(nreg - (q + 1) - 1, outs.data() + (size_t)q * N2);
}

}  // namespace bf
