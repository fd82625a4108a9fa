// bf_device.h -- device-side building blocks of the gfx950 engine.
//
//  * cplx<T>: interleaved complex (T = double for the FP64 path, float for the
//    mixed-precision path), binary compatible with Julia's Complex{T}.
//  * Julia-semantics complex sqrt/exp in double, so that the angular-spectrum
//    phase exp(-i kz dz), kz = sqrt(k0^2 - kx^2 - ky^2), takes the same branch
//    as the reference (real eps carries a -0.0 imaginary part, so evanescent
//    components decay; src/boost_fractor.jl:138-141,171-174).
//  * Line FFTs staged in LDS: Stockham autosort radix-4 passes (+ one radix-2
//    pass when log2 N is odd) for N = 2^k, N/4 threads per line; a direct DFT
//    for any other N (the reference grids 101^2, 51^2, 87^2).  Twiddles come
//    from a per-N table w[m] = exp(-2 pi i m/N) built in FP64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

template <typename T>
struct alignas(2 * sizeof(T)) cplx {
    T re, im;
};

template <typename T>
__host__ __device__ __forceinline__ cplx<T> cmk(T a, T b) {
    cplx<T> z;
    z.re = a;
    z.im = b;
    return z;
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) {
    return cmk<T>(a.re + b.re, a.im + b.im);
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) {
    return cmk<T>(a.re - b.re, a.im - b.im);
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
    return cmk<T>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
// acc + a b as four fused multiply-adds (real part: acc.re + a.re b.re, then - a.im b.im)
template <typename T>
__host__ __device__ __forceinline__ cplx<T> cmac(cplx<T> acc, cplx<T> a, cplx<T> b) {
    return cmk<T>(fma(-a.im, b.im, fma(a.re, b.re, acc.re)), fma(a.im, b.re, fma(a.re, b.im, acc.im)));
}
// a * conj(b)
template <typename T>
__host__ __device__ __forceinline__ cplx<T> cmulc(cplx<T> a, cplx<T> b) {
    return cmk<T>(a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im);
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> rmul(T s, cplx<T> b) {
    return cmk<T>(s * b.re, s * b.im);
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> cconj(cplx<T> a) {
    return cmk<T>(a.re, -a.im);
}
template <typename To, typename Ti>
__host__ __device__ __forceinline__ cplx<To> ccast(cplx<Ti> a) {
    return cmk<To>((To)a.re, (To)a.im);
}

typedef cplx<double> zd;

// precision tag: FP32 storage with FP32 arithmetic
namespace bf {
struct f32pure {};
}  // namespace bf

// (-1im) * z evaluated as Julia does: Complex(0,-1) * z
__host__ __device__ __forceinline__ zd neg_i(zd z) {
#pragma clang fp contract(off)
    return cmk<double>(0.0 * z.re - (-1.0) * z.im, 0.0 * z.im + (-1.0) * z.re);
}

// Julia Base sqrt(::Complex{Float64}): principal branch, imaginary sign taken
// from the (possibly signed-zero) imaginary input.
__host__ __device__ inline zd csqrt_j(zd z) {
#pragma clang fp contract(off)
    double x = z.re, y = z.im;
    if (x == 0.0 && y == 0.0) return cmk<double>(0.0, y);
    if (__builtin_isinf(y)) return cmk<double>(INFINITY, y);
    if (__builtin_isnan(x) || __builtin_isnan(y)) return cmk<double>(NAN, NAN);
    double rho = sqrt((hypot(x, y) + fabs(x)) * 0.5);
    double xi = rho, eta = y;
    if (rho != 0.0) {
        if (__builtin_isfinite(eta)) eta = (eta / rho) * 0.5;
        if (x < 0.0) {
            xi = fabs(eta);
            eta = copysign(rho, y);
        }
    }
    return cmk<double>(xi, eta);
}

// Julia Base exp(::Complex{Float64})
__host__ __device__ inline zd cexp_j(zd z) {
    double er = exp(z.re);
    if (z.im == 0.0) return cmk<double>(er, z.im);
    double s, c;
#ifdef __HIP_DEVICE_COMPILE__
    sincos(z.im, &s, &c);
#else
    s = sin(z.im);
    c = cos(z.im);
#endif
    return cmk<double>(er * c, er * s);
}

// Angular-spectrum phase exp(-1im*kz*dz) with kz = sqrt(k0^2 - Kx^2 - Ky^2),
// op order of src/boost_fractor.jl:171-174.
// kz = sqrt(k0^2 - Kx^2 - Ky^2)
__host__ __device__ inline zd kz_of(zd k0sq, double Kx, double Ky) {
#pragma clang fp contract(off)
    zd arg = cmk<double>(k0sq.re - Kx * Kx, k0sq.im);
    arg = cmk<double>(arg.re - Ky * Ky, arg.im);
    return csqrt_j(arg);
}
// exp(-1im*kz*dz)
__host__ __device__ inline zd kz_phase_of(zd kz, double dz) {
#pragma clang fp contract(off)
    zd e = neg_i(kz);
    return cexp_j(cmk<double>(dz * e.re, dz * e.im));
}
__host__ __device__ inline zd kz_phase(zd k0sq, double Kx, double Ky, double dz) {
    return kz_phase_of(kz_of(k0sq, Kx, Ky), dz);
}

// Aperture predicate of the propagators, abs(x^2 + y^2) < diskR^2
// (src/boost_fractor.jl:159, :198), with separately rounded products.
__host__ __device__ inline bool in_aperture(double x, double y, double R2) {
#pragma clang fp contract(off)
    return fabs(x * x + y * y) < R2;
}

// ---------------------------------------------------------------------------
// Raw buffer access.  A buffer load whose byte offset lies beyond the
// resource's size returns zero without touching memory, and such a store is
// dropped; predicated loads and stores therefore need no branches, and every
// load of a pass can be in flight at once.  Offsets are 32-bit byte offsets
// (a resource covers at most one field or table, < 2 GiB).
// ---------------------------------------------------------------------------
constexpr uint32_t BUF_OOB = 0x80000000u;  // out of range for every resource used here

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, uint32_t bytes) {
    // word 3 = 0x00020000: 32-bit data format, the raw-buffer setting of the gfx9 family
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}

template <typename S>
__device__ __forceinline__ cplx<S> buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off);
template <>
__device__ __forceinline__ cplx<float> buf_load<float>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0);
    return cmk<float>(__uint_as_float(v[0]), __uint_as_float(v[1]));
}
template <>
__device__ __forceinline__ cplx<double> buf_load<double>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
    return cmk<double>(__hiloint2double((int)v[1], (int)v[0]), __hiloint2double((int)v[3], (int)v[2]));
}

template <typename S>
__device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r, uint32_t off, cplx<S> a);
template <>
__device__ __forceinline__ void buf_store<float>(__amdgpu_buffer_rsrc_t r, uint32_t off, cplx<float> a) {
    unsigned int v __attribute__((__vector_size__(8)));
    v[0] = __float_as_uint(a.re);
    v[1] = __float_as_uint(a.im);
    __builtin_amdgcn_raw_buffer_store_b64(v, r, (int)off, 0, 0);
}
template <>
__device__ __forceinline__ void buf_store<double>(__amdgpu_buffer_rsrc_t r, uint32_t off, cplx<double> a) {
    unsigned int v __attribute__((__vector_size__(16)));
    const unsigned long long re = (unsigned long long)__double_as_longlong(a.re);
    const unsigned long long im = (unsigned long long)__double_as_longlong(a.im);
    v[0] = (unsigned int)re;
    v[1] = (unsigned int)(re >> 32);
    v[2] = (unsigned int)im;
    v[3] = (unsigned int)(im >> 32);
    __builtin_amdgcn_raw_buffer_store_b128(v, r, (int)off, 0, 0);
}

// ---------------------------------------------------------------------------
// Line FFTs in LDS
// ---------------------------------------------------------------------------
// threads per line for a compile-time length N (0 = runtime length, DFT path)
template <int N>
struct LineCfg {
    static constexpr int TPL = (N >= 16) ? N / 4 : 64;
};

template <typename T, bool INV>
__device__ __forceinline__ cplx<T> tw_get(const cplx<T>* __restrict__ tw, int m) {
    cplx<T> w = tw[m];
    if (INV) w.im = -w.im;
    return w;
}

// radix-4 DFT of (a0..a3) in place; forward uses w = -i, inverse w = +i
template <typename T, bool INV>
__device__ __forceinline__ void radix4(cplx<T>& a0, cplx<T>& a1, cplx<T>& a2, cplx<T>& a3) {
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

// Stockham autosort FFT of one line of N = 2^k points held in LDS `s`.
// Thread t in [0, N/4).  All threads of the block must call it (barriers).
template <typename T, int N, bool INV>
__device__ __forceinline__ void fft_pow2(cplx<T>* s, int t, const cplx<T>* __restrict__ tw) {
    constexpr int Q = N / 4;
#pragma unroll
    for (int Ns = 1; Ns * 4 <= N; Ns *= 4) {
        cplx<T> a0 = s[t], a1 = s[t + Q], a2 = s[t + 2 * Q], a3 = s[t + 3 * Q];
        const int k = t & (Ns - 1);
        if (Ns > 1) {
            const int st = N / (4 * Ns);
            a1 = cmul(a1, tw_get<T, INV>(tw, k * st));
            a2 = cmul(a2, tw_get<T, INV>(tw, 2 * k * st));
            a3 = cmul(a3, tw_get<T, INV>(tw, 3 * k * st));
        }
        radix4<T, INV>(a0, a1, a2, a3);
        __syncthreads();
        const int base = ((t - k) << 2) + k;
        s[base] = a0;
        s[base + Ns] = a1;
        s[base + 2 * Ns] = a2;
        s[base + 3 * Ns] = a3;
        __syncthreads();
    }
    constexpr bool odd = (__builtin_ctz(N) & 1) != 0;
    if (odd) {  // final radix-2 pass, Ns = N/2: thread t owns j = t and j = t + Q
        constexpr int H = N / 2;
        cplx<T> a = s[t], b = s[t + H], c = s[t + Q], d = s[t + Q + H];
        b = cmul(b, tw_get<T, INV>(tw, t));
        d = cmul(d, tw_get<T, INV>(tw, t + Q));
        s[t] = cadd(a, b);
        s[t + H] = csub(a, b);
        s[t + Q] = cadd(c, d);
        s[t + Q + H] = csub(c, d);
        __syncthreads();
    }
}

// Direct DFT of one line of n points: in -> out (distinct LDS buffers), thread
// t of nt handles outputs k = t, t+nt, ...  Caller syncs before reading `out`.
template <typename T, bool INV>
__device__ __forceinline__ void dft_line(const cplx<T>* in, cplx<T>* out, int n, int t, int nt,
                                         const cplx<T>* __restrict__ tw) {
    for (int k = t; k < n; k += nt) {
        cplx<T> acc = cmk<T>(0, 0);
        int idx = 0;
        for (int j = 0; j < n; j++) {
            acc = cadd(acc, cmul(in[j], tw_get<T, INV>(tw, idx)));
            idx += k;
            if (idx >= n) idx -= n;
        }
        out[k] = acc;
    }
}

// Transform one line in place in `s` (length n; N compile-time or 0).
// `s2` is a second LDS line buffer used by the DFT path.
template <typename T, int N, bool INV>
__device__ __forceinline__ void line_fft(cplx<T>* s, cplx<T>* s2, int n, int t, const cplx<T>* __restrict__ tw) {
    if constexpr (N >= 16) {
        fft_pow2<T, N, INV>(s, t, tw);
    } else {
        constexpr int TPL = LineCfg<N>::TPL;
        if (n > 1) dft_line<T, INV>(s, s2, n, t, TPL, tw);
        __syncthreads();
        if (n > 1)
            for (int k = t; k < n; k += TPL) s[k] = s2[k];
        __syncthreads();
    }
}
