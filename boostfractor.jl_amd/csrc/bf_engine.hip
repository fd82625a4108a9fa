// bf_engine.hip -- recursive Fourier propagation on gfx950.
//
// Data layout in HBM: every field is a dense [ny][nx] complex array (x fastest,
// Julia's E[x, y]); a batch of frequencies keeps its fields in one buffer
// [freq][region][slot][ny][nx].  A 2D propagation (src/boost_fractor.jl:157-179)
// is split into passes that each stream every field through LDS once:
//
//   row pass    (per line y):   [row IFFT] -> [x tilt*surface] -> [mask] -> [row FFT]
//   column pass (per x tile):   [col FFT] -> x exp(-i kz dz)/K -> [col IFFT]
//
// The dancer bounce (src/boost_fractor_recurprop.jl:71-133) fuses the
// boundary update into the row pass: a block loads row y of the two fields that
// meet at boundary b, inverse-transforms both, applies tilt/surface, the
// reflection/transmission update (and the Eout accumulation at the exit
// boundary), masks and forward-transforms both for the next bounce.  A bounce
// is therefore exactly two launches (column pass, then fused row pass), each
// reading and writing every field once.  The per-region k-space phase tables
// are built once per frequency batch from FP64 arguments.
#include "bf_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace bf {

// ---------------------------------------------------------------------------
// descriptors
// ---------------------------------------------------------------------------
enum : uint32_t { RF_INV = 1, RF_RTAB = 2, RF_MASK = 4, RF_FWD = 8 };
enum : uint32_t { CF_FWD = 1, CF_KMID = 2, CF_INV = 4, CF_KLOAD = 8, CF_INSHIFT = 16, CF_OUTSHIFT = 32 };
enum : uint32_t { DF_FWD = 1, DF_AMAX = 2, DF_EOUT_ASSIGN = 4 };

// ---------------------------------------------------------------------------
// small kernels
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_twiddles(cplx<T>* out, int n) {
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    double s, c;
    sincospi(2.0 * (double)m / (double)n, &s, &c);
    out[m] = cmk<T>((T)c, (T)(-s));
}

template <typename T>
__global__ void k_build_ktab(cplx<T>* out, const KTabDesc* __restrict__ d, int nx, int ny,
                             const double* __restrict__ kX, const double* __restrict__ kY, double scale) {
    const int tab = blockIdx.y;
    const size_t K = (size_t)nx * ny;
    const KTabDesc dd = d[tab];
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < K; i += (size_t)gridDim.x * blockDim.x) {
        int mx = (int)(i % nx), my = (int)(i / nx);
        int jx = mx, jy = my;
        if (!dd.shifted) {  // unshifted bin m multiplies the fftshifted entry (m + floor(N/2)) mod N
            jx = (mx + nx / 2) % nx;
            jy = (my + ny / 2) % ny;
        }
        zd ph = kz_phase(dd.k0sq, kX[jx], kY[jy], dd.dz);
        out[(size_t)tab * K + i] = cmk<T>((T)(ph.re * scale), (T)(ph.im * scale));
    }
}

template <typename T>
__global__ void k_build_rtab(cplx<T>* out, const RTabDesc* __restrict__ d, int nx, int ny,
                             const double* __restrict__ X, const double* __restrict__ Y, const zd* __restrict__ surf) {
    const int tab = blockIdx.y;
    const size_t K = (size_t)nx * ny;
    const RTabDesc dd = d[tab];
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < K; i += (size_t)gridDim.x * blockDim.x) {
        int x = (int)(i % nx), y = (int)(i / nx);
        // exp(-1im*k0*tilt_x*x) * exp(-1im*k0*tilt_y*y)   (src/boost_fractor.jl:145)
        zd px = cexp_j(rmul(X[x], rmul(dd.tx, dd.a)));
        zd py = cexp_j(rmul(Y[y], rmul(dd.ty, dd.a)));
        zd v = cmul(px, py);
        if (dd.surf >= 0) v = cmul(v, cexp_j(cmul(dd.a, surf[dd.surf + i])));  // exp(-1im*k0*surface) (:147)
        out[(size_t)tab * K + i] = cmk<T>((T)v.re, (T)v.im);
    }
}

// fields[j] *= ph[j] (propagator1D, src/boost_fractor.jl:218-227)
template <typename T>
__global__ void k_scale_fields(cplx<T>* base, const FieldDesc* __restrict__ fd, const zd* __restrict__ ph, size_t K,
                               const int* __restrict__ active) {
    const int j = blockIdx.y;
    const FieldDesc d = fd[j];
    if (active && !active[d.group]) return;
    const cplx<T> p = cmk<T>((T)ph[j].re, (T)ph[j].im);
    cplx<T>* f = base + d.off;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < K; i += (size_t)gridDim.x * blockDim.x)
        f[i] = cmul(f[i], p);
}

// ---------------------------------------------------------------------------
// row pass
// ---------------------------------------------------------------------------
template <typename T>
struct RowArgs {
    cplx<T>* base;
    const FieldDesc* fd;
    int nf, nx, ny;
    const cplx<T>* tw;
    uint32_t flags;
    const cplx<T>* rtab;
    const double* X;
    const double* Y;
    double R2;
    const int* active;
};

template <typename T, int N>
__global__ void __launch_bounds__(1024) k_rows(RowArgs<T> a) {
    extern __shared__ unsigned char smem_raw[];
    cplx<T>* smem = reinterpret_cast<cplx<T>*>(smem_raw);
    constexpr int TPL = LineCfg<N>::TPL;
    const int n = (N >= 16) ? N : a.nx;
    const int lpb = blockDim.x / TPL;
    const int li = threadIdx.x / TPL, t = threadIdx.x % TPL;
    cplx<T>* s = smem + (size_t)li * n;
    cplx<T>* s2 = smem + (size_t)(lpb + li) * n;
    const long nlines = (long)a.nf * a.ny;
    const long line = (long)blockIdx.x * lpb + li;
    bool valid = line < nlines;
    const int j = valid ? (int)(line / a.ny) : 0;
    const int y = valid ? (int)(line % a.ny) : 0;
    const FieldDesc d = a.fd[j];
    if (valid && a.active && !a.active[d.group]) valid = false;
    cplx<T>* row = a.base + d.off + (size_t)y * a.nx;
    for (int x = t; x < n; x += TPL) s[x] = valid ? row[x] : cmk<T>(0, 0);
    __syncthreads();
    if (a.flags & RF_INV) line_fft<T, N, true>(s, s2, n, t, a.tw);
    if (a.flags & (RF_RTAB | RF_MASK)) {
        const cplx<T>* tab = ((a.flags & RF_RTAB) && d.rtab >= 0)
                                 ? a.rtab + ((size_t)d.rtab * a.ny + y) * a.nx
                                 : nullptr;
        const double yy = a.Y[y];
        for (int x = t; x < n; x += TPL) {
            cplx<T> v = s[x];
            if (tab) v = cmul(v, tab[x]);
            if (a.flags & RF_MASK) {
                const double xx = a.X[x];
                if (!(fabs(xx * xx + yy * yy) < a.R2)) v = cmk<T>(0, 0);
            }
            s[x] = v;
        }
        __syncthreads();
    }
    if (a.flags & RF_FWD) line_fft<T, N, false>(s, s2, n, t, a.tw);
    if (valid)
        for (int x = t; x < n; x += TPL) row[x] = s[x];
}

// ---------------------------------------------------------------------------
// column pass
// ---------------------------------------------------------------------------
template <typename T>
struct ColArgs {
    const cplx<T>* src;
    cplx<T>* dst;
    const FieldDesc* fd;
    int nf, nx, ny;
    const cplx<T>* tw;
    uint32_t flags;
    const cplx<T>* ktab;
    int tx;  // columns per block
    const int* active;
};

template <typename T, int N>
__global__ void __launch_bounds__(1024) k_cols(ColArgs<T> a) {
    extern __shared__ unsigned char smem_raw[];
    cplx<T>* smem = reinterpret_cast<cplx<T>*>(smem_raw);
    constexpr int TPL = LineCfg<N>::TPL;
    const int n = (N >= 16) ? N : a.ny;
    const int ld = n + 1;
    const int tx = a.tx;
    cplx<T>* s = smem;
    cplx<T>* s2 = smem + (size_t)tx * ld;
    const FieldDesc d = a.fd[blockIdx.y];
    if (a.active && !a.active[d.group]) return;  // uniform over the block
    const int x0 = blockIdx.x * tx;
    const int nx = a.nx, ny = a.ny;
    const int hx = nx / 2, hy = ny / 2;
    const cplx<T>* src = a.src + d.off;
    cplx<T>* dst = a.dst + d.off;
    const cplx<T>* kt = (d.ktab >= 0) ? a.ktab + (size_t)d.ktab * ny * nx : nullptr;
    const uint32_t fl = a.flags;
    for (int i = threadIdx.x; i < tx * n; i += blockDim.x) {
        const int c = i % tx, my = i / tx, mx = x0 + c;
        cplx<T> v = cmk<T>(0, 0);
        if (mx < nx) {
            int sx = mx, sy = my;
            if (fl & CF_INSHIFT) {  // ifftshift on load
                sx = (mx + hx) % nx;
                sy = (my + hy) % ny;
            }
            const size_t o = (size_t)sy * nx + sx;
            v = src[o];
            if ((fl & CF_KLOAD) && kt) v = cmul(v, kt[o]);
        }
        s[c * ld + my] = v;
    }
    __syncthreads();
    const int c = threadIdx.x / TPL, t = threadIdx.x % TPL;
    cplx<T>* line = s + c * ld;
    cplx<T>* line2 = s2 + c * ld;
    if (fl & CF_FWD) line_fft<T, N, false>(line, line2, n, t, a.tw);
    if ((fl & CF_KMID) && kt) {
        for (int i = threadIdx.x; i < tx * n; i += blockDim.x) {
            const int cc = i % tx, my = i / tx, mx = x0 + cc;
            if (mx < nx) s[cc * ld + my] = cmul(s[cc * ld + my], kt[(size_t)my * nx + mx]);
        }
        __syncthreads();
    }
    if (fl & CF_INV) line_fft<T, N, true>(line, line2, n, t, a.tw);
    for (int i = threadIdx.x; i < tx * n; i += blockDim.x) {
        const int cc = i % tx, my = i / tx, mx = x0 + cc;
        if (mx < nx) {
            int dx = mx, dy = my;
            if (fl & CF_OUTSHIFT) {  // fftshift on store
                dx = (mx + hx) % nx;
                dy = (my + hy) % ny;
            }
            dst[(size_t)dy * nx + dx] = s[cc * ld + my];
        }
    }
}

// ---------------------------------------------------------------------------
// dancer: fused row pass with the boundary update
// ---------------------------------------------------------------------------
template <typename T>
struct DancerRowArgs {
    cplx<T>* fields;  // [fb][R][2][ny][nx]
    int R, nx, ny;
    int rm;  // slot holding the right-moving fields this bounce
    const cplx<T>* tw;
    const cplx<T>* rtab;
    const int* rtab_idx;  // [fb][R], -1: none
    const zd* coef;       // boundary reflectivities r_b, b = 0..R (r_R = 0 appended)
    const double* X;
    const double* Y;
    double R2;
    uint32_t flags;
    zd* Eout;  // [nslices][ny][nx] per frequency slot, stride eout_fstride
    size_t eout_fstride;
    int slice;
    double* amax;  // [fb]
    const int* active;
    int* count;  // [fb] bounce counters
};

__device__ __forceinline__ void atomic_max_pos(double* addr, double v) {
    // non-negative doubles order like their bit patterns
    atomicMax(reinterpret_cast<unsigned long long*>(addr), (unsigned long long)__double_as_longlong(v));
}

template <typename T, int N>
__global__ void __launch_bounds__(1024) k_dancer_rows(DancerRowArgs<T> a) {
    extern __shared__ unsigned char smem_raw[];
    cplx<T>* smem = reinterpret_cast<cplx<T>*>(smem_raw);
    __shared__ double red[32];
    constexpr int TPL = LineCfg<N>::TPL;
    const int n = (N >= 16) ? N : a.nx;
    const int f = blockIdx.z, b = blockIdx.y;
    if (a.active && !a.active[f]) return;  // uniform over the block
    if (a.count && blockIdx.x == 0 && b == 0 && threadIdx.x == 0) a.count[f] += 1;
    const int lpb = blockDim.x / TPL;  // 2 * rows per block
    const int li = threadIdx.x / TPL, t = threadIdx.x % TPL;
    const int side = li & 1;  // 0: A = region b-1 right-moving, 1: B = region b left-moving
    const int y = blockIdx.x * (lpb >> 1) + (li >> 1);
    const int reg = side == 0 ? b - 1 : b;
    const int slot = side == 0 ? a.rm : 1 - a.rm;
    const bool exists = reg >= 0 && reg < a.R && y < a.ny;
    const size_t K = (size_t)a.nx * a.ny;
    cplx<T>* s = smem + (size_t)li * n;
    cplx<T>* s2 = smem + (size_t)(lpb + li) * n;
    cplx<T>* row = a.fields + (((size_t)f * a.R + (exists ? reg : 0)) * 2 + slot) * K + (size_t)(exists ? y : 0) * a.nx;
    for (int x = t; x < n; x += TPL) s[x] = exists ? row[x] : cmk<T>(0, 0);
    __syncthreads();
    line_fft<T, N, true>(s, s2, n, t, a.tw);
    if (a.rtab) {
        const int ti = exists ? a.rtab_idx[f * a.R + reg] : -1;
        if (ti >= 0) {
            const cplx<T>* tab = a.rtab + (size_t)ti * K + (size_t)y * a.nx;
            for (int x = t; x < n; x += TPL) s[x] = cmul(s[x], tab[x]);
        }
        __syncthreads();
    }
    // boundary update (src/boost_fractor_recurprop.jl:90-119), result staged in s2
    const zd r = a.coef[b];
    const cplx<T> rT = cmk<T>((T)r.re, (T)r.im);
    const cplx<T> mrT = cmk<T>((T)(-r.re), (T)(-r.im));
    const cplx<T> omrT = cmk<T>((T)(1.0 - r.re), (T)(-r.im));
    const cplx<T> oprT = cmk<T>((T)(1.0 + r.re), (T)r.im);
    const cplx<T>* p = smem + (size_t)(li ^ 1) * n;  // partner line, same y
    double vmax = 0.0;
    const double yy = a.Y[exists ? y : 0];
    for (int x = t; x < n; x += TPL) {
        const cplx<T> v = s[x], w = p[x];
        cplx<T> nv = cmk<T>(0, 0);
        if (exists) {
            if (side == 0) {
                if (b == a.R) {  // exit boundary: Eout += (1+r)*A, A *= r
                    const zd A = ccast<double>(v);
                    const zd e = cmul(cmk<double>(1.0 + r.re, r.im), A);
                    zd* eo = a.Eout + (size_t)f * a.eout_fstride + (size_t)a.slice * K + (size_t)y * a.nx + x;
                    *eo = (a.flags & DF_EOUT_ASSIGN) ? e : cadd(*eo, e);
                    nv = cmul(v, rT);
                } else {  // A' = r A + (1-r) B
                    nv = cadd(cmul(v, rT), cmul(omrT, w));
                }
            } else {
                if (b == 0)  // mirror: B' = -r B
                    nv = cmul(v, mrT);
                else  // B' = -r B + (1+r) A
                    nv = cadd(cmul(v, mrT), cmul(oprT, w));
            }
            if (a.flags & DF_AMAX) vmax = fmax(vmax, hypot((double)nv.re, (double)nv.im));
            if (a.flags & DF_FWD) {
                const double xx = a.X[x];
                if (!(fabs(xx * xx + yy * yy) < a.R2)) nv = cmk<T>(0, 0);  // next bounce's aperture
            }
        }
        s2[x] = nv;
    }
    if (a.flags & DF_AMAX) {
        for (int o = 32; o > 0; o >>= 1) vmax = fmax(vmax, __shfl_xor(vmax, o));
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = vmax;
    }
    __syncthreads();
    if ((a.flags & DF_AMAX) && threadIdx.x == 0) {
        double m = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) m = fmax(m, red[w]);
        if (m > 0.0) atomic_max_pos(a.amax + f, m);
    }
    if (!(a.flags & DF_FWD)) return;  // last bounce: only Eout is needed
    for (int x = t; x < n; x += TPL) s[x] = s2[x];
    __syncthreads();
    line_fft<T, N, false>(s, s2, n, t, a.tw);
    if (exists)
        for (int x = t; x < n; x += TPL) row[x] = s[x];
}

// pointwise bounce update for the unfused propagators (propagator1D, momentum space)
template <typename T>
__global__ void k_dancer_update(cplx<T>* fields, int R, size_t K, int rm, const zd* __restrict__ coef, zd* Eout,
                                size_t eout_fstride, int slice, int assign, double* amax, const int* active,
                                int* count) {
    const int f = blockIdx.z, b = blockIdx.y;
    if (active && !active[f]) return;
    if (count && blockIdx.x == 0 && b == 0 && threadIdx.x == 0) count[f] += 1;
    const zd r = coef[b];
    const cplx<T> rT = cmk<T>((T)r.re, (T)r.im), mrT = cmk<T>((T)(-r.re), (T)(-r.im));
    const cplx<T> omrT = cmk<T>((T)(1.0 - r.re), (T)(-r.im)), oprT = cmk<T>((T)(1.0 + r.re), (T)r.im);
    cplx<T>* A = (b >= 1) ? fields + (((size_t)f * R + (b - 1)) * 2 + rm) * K : nullptr;
    cplx<T>* B = (b < R) ? fields + (((size_t)f * R + b) * 2 + (1 - rm)) * K : nullptr;
    double vmax = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < K; i += (size_t)gridDim.x * blockDim.x) {
        if (b == 0) {
            B[i] = cmul(B[i], mrT);
            vmax = fmax(vmax, hypot((double)B[i].re, (double)B[i].im));
        } else if (b == R) {
            const zd e = cmul(cmk<double>(1.0 + r.re, r.im), ccast<double>(A[i]));
            zd* eo = Eout + (size_t)f * eout_fstride + (size_t)slice * K + i;
            *eo = assign ? e : cadd(*eo, e);
            A[i] = cmul(A[i], rT);
            vmax = fmax(vmax, hypot((double)A[i].re, (double)A[i].im));
        } else {
            const cplx<T> va = A[i], vb = B[i];
            const cplx<T> na = cadd(cmul(va, rT), cmul(omrT, vb));
            const cplx<T> nb = cadd(cmul(vb, mrT), cmul(oprT, va));
            A[i] = na;
            B[i] = nb;
            vmax = fmax(vmax, fmax(hypot((double)na.re, (double)na.im), hypot((double)nb.re, (double)nb.im)));
        }
    }
    if (amax) {
        for (int o = 32; o > 0; o >>= 1) vmax = fmax(vmax, __shfl_xor(vmax, o));
        if ((threadIdx.x & 63) == 0 && vmax > 0.0) atomic_max_pos(amax + f, vmax);
    }
}

// while n <= nmax && a >= amin: deactivate frequency slots whose field maximum fell below amin
__global__ void k_amin_check(int* active, double* amax, double amin, int fb) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= fb) return;
    if (active[f] && !(amax[f] >= amin)) active[f] = 0;
    amax[f] = 0.0;
}

// dance_intro fill (src/boost_fractor_recurprop.jl:153-201): fields[f][n][s] = ax[n][s] * [x^2+y^2 <= R^2]
template <typename T>
__global__ void k_dance_intro(cplx<T>* fields, int nd, size_t K, int nx, const zd* __restrict__ ax,
                              const double* __restrict__ X, const double* __restrict__ Y, double R2,
                              size_t fstride) {
    const int f = blockIdx.z, q = blockIdx.y;  // q = region*2 + slot
    const zd v = ax[q];
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < K; i += (size_t)gridDim.x * blockDim.x) {
        const double xx = X[i % nx], yy = Y[i / nx];
        const double bb = (xx * xx + yy * yy > R2) ? 0.0 : 1.0;
        fields[(size_t)f * fstride + (size_t)q * K + i] = cmk<T>((T)(bb * v.re), (T)(bb * v.im));
    }
}

// broadcast a double field block [nq][K] into every frequency slot of a T buffer
template <typename T>
__global__ void k_broadcast(cplx<T>* dst, const zd* __restrict__ src, size_t n, size_t fstride) {
    const int f = blockIdx.y;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[(size_t)f * fstride + i] = cmk<T>((T)src[i].re, (T)src[i].im);
}

template <typename T>
__global__ void k_to_double(zd* dst, const cplx<T>* __restrict__ src, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = ccast<double>(src[i]);
}

// sum(reverse(Eout, dims=3), dims=3): out[f] = ((E[S-1] + E[S-2]) + ...) + E[0]
__global__ void k_reverse_sum(zd* out, const zd* __restrict__ hist, int S, size_t K) {
    const int f = blockIdx.y;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < K; i += (size_t)gridDim.x * blockDim.x) {
        zd s = cmk<double>(0.0, 0.0);
        for (int q = S - 1; q >= 0; q--) s = cadd(s, hist[((size_t)f * S + q) * K + i]);
        out[(size_t)f * K + i] = s;
    }
}

// power[f] = sum |E_f|^2 (two-level: per-block partials then one block)
__global__ void k_power(const zd* __restrict__ E, size_t K, double* out) {
    const int f = blockIdx.x;
    __shared__ double red[16];
    double s = 0.0;
    for (size_t i = threadIdx.x; i < K; i += blockDim.x) {
        const zd v = E[(size_t)f * K + i];
        s += v.re * v.re + v.im * v.im;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += red[w];
        out[f] = t;
    }
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
void allow_lds(const void* kernel, size_t bytes) {
    if (bytes <= 64 * 1024) return;
    BF_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

void upload(void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (bytes) BF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
}
void download(void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (bytes) BF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
}

void Grid::upload(const bf_coords* c, hipStream_t s) {
    nx = c->nx;
    ny = c->ny;
    X.alloc(sizeof(double) * nx);
    Y.alloc(sizeof(double) * ny);
    kX.alloc(sizeof(double) * nx);
    kY.alloc(sizeof(double) * ny);
    bf::upload(X.p, c->X, sizeof(double) * nx, s);
    bf::upload(Y.p, c->Y, sizeof(double) * ny, s);
    bf::upload(kX.p, c->kX, sizeof(double) * nx, s);
    bf::upload(kY.p, c->kY, sizeof(double) * ny, s);
}

template <typename T>
const cplx<T>* twiddle_table(bf_ctx* ctx, int n) {
    auto key = std::make_pair(n, (int)sizeof(T));
    auto it = ctx->twiddles.find(key);
    if (it != ctx->twiddles.end()) return it->second.as<cplx<T>>();
    DevBuf b(sizeof(cplx<T>) * (size_t)std::max(n, 1));
    hipLaunchKernelGGL(k_twiddles<T>, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, b.as<cplx<T>>(), n);
    BF_CHECK_LAUNCH();
    const cplx<T>* p = b.as<cplx<T>>();
    ctx->twiddles.emplace(key, std::move(b));
    return p;
}
template const cplx<double>* twiddle_table<double>(bf_ctx*, int);
template const cplx<float>* twiddle_table<float>(bf_ctx*, int);

static inline unsigned grid1d(size_t n, int threads, unsigned cap = 2048) {
    size_t b = (n + threads - 1) / threads;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

// compile-time line length dispatch: 16..1024 (powers of two) or 0 (DFT)
template <template <typename, int> class F, typename T, typename... A>
static void dispatch_n(int n, A&&... args) {
    switch (n) {
        case 16: F<T, 16>::run(std::forward<A>(args)...); break;
        case 32: F<T, 32>::run(std::forward<A>(args)...); break;
        case 64: F<T, 64>::run(std::forward<A>(args)...); break;
        case 128: F<T, 128>::run(std::forward<A>(args)...); break;
        case 256: F<T, 256>::run(std::forward<A>(args)...); break;
        case 512: F<T, 512>::run(std::forward<A>(args)...); break;
        case 1024: F<T, 1024>::run(std::forward<A>(args)...); break;
        default: F<T, 0>::run(std::forward<A>(args)...); break;
    }
}

template <int N>
static int tpl_of() { return LineCfg<N>::TPL; }

template <typename T, int N>
struct RowLaunch {
    static void run(const RowArgs<T>& a, hipStream_t s) {
        const int TPL = LineCfg<N>::TPL;
        const int n = N >= 16 ? N : a.nx;
        const int lpb = std::max(1, 256 / TPL);
        const long nlines = (long)a.nf * a.ny;
        const size_t smem = sizeof(cplx<T>) * (size_t)n * lpb * (N >= 16 ? 1 : 2);
        allow_lds((const void*)k_rows<T, N>, smem);
        hipLaunchKernelGGL((k_rows<T, N>), dim3((unsigned)((nlines + lpb - 1) / lpb)), dim3(lpb * TPL), smem, s, a);
        BF_CHECK_LAUNCH();
    }
};

template <typename T, int N>
struct ColLaunch {
    static void run(ColArgs<T> a, int tx_req, hipStream_t s) {
        const int TPL = LineCfg<N>::TPL;
        const int n = N >= 16 ? N : a.ny;
        int tx = tx_req > 0 ? tx_req : (int)(128 / sizeof(cplx<T>));
        tx = std::max(1, std::min(tx, 1024 / TPL));
        while (tx > 1 && sizeof(cplx<T>) * (size_t)tx * (n + 1) * (N >= 16 ? 1 : 2) > 96 * 1024) tx >>= 1;
        a.tx = tx;
        const size_t smem = sizeof(cplx<T>) * (size_t)tx * (n + 1) * (N >= 16 ? 1 : 2);
        allow_lds((const void*)k_cols<T, N>, smem);
        hipLaunchKernelGGL((k_cols<T, N>), dim3((a.nx + tx - 1) / tx, a.nf), dim3(tx * TPL), smem, s, a);
        BF_CHECK_LAUNCH();
    }
};

template <typename T, int N>
struct DancerRowLaunch {
    static void run(const DancerRowArgs<T>& a, int fb, int rows_req, hipStream_t s) {
        const int TPL = LineCfg<N>::TPL;
        const int n = N >= 16 ? N : a.nx;
        int rows = rows_req > 0 ? rows_req : std::max(1, 128 / TPL);
        rows = std::max(1, std::min(rows, 512 / TPL));
        const size_t smem = sizeof(cplx<T>) * (size_t)n * (2 * rows) * 2;
        allow_lds((const void*)k_dancer_rows<T, N>, smem);
        hipLaunchKernelGGL((k_dancer_rows<T, N>), dim3((a.ny + rows - 1) / rows, a.R + 1, fb), dim3(2 * rows * TPL),
                           smem, s, a);
        BF_CHECK_LAUNCH();
    }
};

// ---------------------------------------------------------------------------
// Propagator engine over a set of fields
// ---------------------------------------------------------------------------
// Julia: k0 = 2*pi/lambda*sqrt(eps)
static inline zd k0_of(zd eps, double lambda) { return rmul((2.0 * M_PI) / lambda, csqrt_j(eps)); }

// One propagation region at one frequency: what a prop() call needs.
struct RegionParams {
    zd eps;
    double dz, tx, ty, lambda;
    int64_t surf;  // element offset into the surfaces buffer or -1 (flat)
};

template <typename T>
struct PropEngine {
    bf_ctx* ctx;
    const Grid* g;
    int kind;
    double R2;
    const cplx<T>* twx;
    const cplx<T>* twy;
    const zd* surfaces;  // device, may be null
    DevBuf ktab, rtab, kdesc, rdesc, scratch;
    int ntab = 0;
    std::vector<int> rtab_of;  // per table: rtab index or -1
    int col_tx = 0;            // autotuned column tile (0: default)

    PropEngine(bf_ctx* c, const Grid* gr, int k, double diskR, const zd* surf)
        : ctx(c), g(gr), kind(k), R2(diskR * diskR), surfaces(surf) {
        twx = twiddle_table<T>(ctx, g->nx);
        twy = twiddle_table<T>(ctx, g->ny);
    }

    // Build the k-space and real-space tables for a list of regions.
    void build_tables(const std::vector<RegionParams>& regs) {
        const size_t K = g->K();
        ntab = (int)regs.size();
        std::vector<KTabDesc> kd(ntab);
        std::vector<RTabDesc> rd;
        rtab_of.assign(ntab, -1);
        for (int i = 0; i < ntab; i++) {
            const RegionParams& p = regs[i];
            zd eps = p.eps;
            if (kind == BF_PROP_PROPAGATOR && eps.im == 0.0) eps.im = -0.0;  // src/boost_fractor.jl:140
            const zd k0 = k0_of(eps, p.lambda);
            kd[i].k0sq = cmul(k0, k0);
            kd[i].dz = p.dz;
            kd[i].shifted = (kind == BF_PROP_MOMENTUM) ? 1 : 0;
            bool need_r = false;
            if (kind == BF_PROP_PROPAGATOR) need_r = (p.tx != 0.0 || p.ty != 0.0 || p.surf >= 0);
            if (kind == BF_PROP_MOMENTUM) need_r = (p.tx != 0.0 || p.ty != 0.0);  // :201 tilt only
            if (need_r) {
                RTabDesc r;
                r.a = neg_i(k0);
                r.tx = p.tx;
                r.ty = p.ty;
                r.surf = (kind == BF_PROP_PROPAGATOR) ? p.surf : -1;
                rtab_of[i] = (int)rd.size();
                rd.push_back(r);
            }
        }
        if (kind == BF_PROP_1D) return;
        const double scale = 1.0 / (double)((int64_t)g->nx * g->ny);  // ifft normalisation folded in
        if (ktab.bytes < sizeof(cplx<T>) * K * ntab) ktab.alloc(sizeof(cplx<T>) * K * ntab);
        if (kdesc.bytes < sizeof(KTabDesc) * ntab) kdesc.alloc(sizeof(KTabDesc) * ntab);
        upload(kdesc.p, kd.data(), sizeof(KTabDesc) * ntab, ctx->stream);
        hipLaunchKernelGGL(k_build_ktab<T>, dim3(grid1d(K, 256, 1024), ntab), dim3(256), 0, ctx->stream,
                           ktab.as<cplx<T>>(), kdesc.as<KTabDesc>(), g->nx, g->ny, g->kX.as<double>(),
                           g->kY.as<double>(), scale);
        BF_CHECK_LAUNCH();
        if (!rd.empty()) {
            if (rtab.bytes < sizeof(cplx<T>) * K * rd.size()) rtab.alloc(sizeof(cplx<T>) * K * rd.size());
            if (rdesc.bytes < sizeof(RTabDesc) * rd.size()) rdesc.alloc(sizeof(RTabDesc) * rd.size());
            upload(rdesc.p, rd.data(), sizeof(RTabDesc) * rd.size(), ctx->stream);
            hipLaunchKernelGGL(k_build_rtab<T>, dim3(grid1d(K, 256, 1024), (unsigned)rd.size()), dim3(256), 0,
                               ctx->stream, rtab.as<cplx<T>>(), rdesc.as<RTabDesc>(), g->nx, g->ny, g->X.as<double>(),
                               g->Y.as<double>(), surfaces);
            BF_CHECK_LAUNCH();
        }
    }

    bool has_rtab() const { return rtab.p != nullptr && !rtab_of.empty() &&
                                   std::any_of(rtab_of.begin(), rtab_of.end(), [](int v) { return v >= 0; }); }

    void rows(cplx<T>* base, const FieldDesc* fd, int nf, uint32_t flags, const int* active) {
        RowArgs<T> a;
        a.base = base; a.fd = fd; a.nf = nf; a.nx = g->nx; a.ny = g->ny; a.tw = twx; a.flags = flags;
        a.rtab = rtab.as<cplx<T>>(); a.X = g->X.as<double>(); a.Y = g->Y.as<double>(); a.R2 = R2; a.active = active;
        dispatch_n<RowLaunch, T>(g->nx, a, ctx->stream);
    }
    void cols(const cplx<T>* src, cplx<T>* dst, const FieldDesc* fd, int nf, uint32_t flags, const int* active) {
        ColArgs<T> a;
        a.src = src; a.dst = dst; a.fd = fd; a.nf = nf; a.nx = g->nx; a.ny = g->ny; a.tw = twy; a.flags = flags;
        a.ktab = ktab.as<cplx<T>>(); a.tx = 0; a.active = active;
        dispatch_n<ColLaunch, T>(g->ny, a, col_tx, ctx->stream);
    }

    // Full propagation of nf fields (field j uses tables fd[j].ktab / fd[j].rtab).
    // `scr` must hold as many elements as `base` spans (momentum-space kind only).
    void propagate(cplx<T>* base, const FieldDesc* fd, int nf, const zd* ph1d, cplx<T>* scr, const int* active) {
        const size_t K = g->K();
        switch (kind) {
            case BF_PROP_PROPAGATOR:
            case BF_PROP_NOTILTS:
                rows(base, fd, nf, RF_MASK | RF_FWD, active);
                cols(base, base, fd, nf, CF_FWD | CF_KMID | CF_INV, active);
                rows(base, fd, nf, RF_INV | (kind == BF_PROP_PROPAGATOR ? RF_RTAB : 0u), active);
                break;
            case BF_PROP_MOMENTUM:
                cols(base, scr, fd, nf, CF_INSHIFT | CF_KLOAD | CF_INV, active);
                rows(scr, fd, nf, RF_INV | RF_MASK | RF_RTAB | RF_FWD, active);
                cols(scr, base, fd, nf, CF_FWD | CF_OUTSHIFT, active);
                break;
            default:
                hipLaunchKernelGGL(k_scale_fields<T>, dim3(grid1d(K, 256, 256), nf), dim3(256), 0, ctx->stream, base,
                                   fd, ph1d, K, active);
                BF_CHECK_LAUNCH();
                break;
        }
    }
};

// scalar phase of propagator1D: exp(-1im*k0*dz), k0 = 2pi/lambda*sqrt(eps) (no -0.0 cast)
static zd phase_1d(zd eps, double lambda, double dz) {
    zd k0 = k0_of(eps, lambda);
    zd e = neg_i(k0);
    return cexp_j(cmk<double>(dz * e.re, dz * e.im));
}

// ---------------------------------------------------------------------------
// single propagation (bf_propagate)
// ---------------------------------------------------------------------------
void propagate_one(bf_ctx* ctx, int kind, const bf_coords* coords, double* E, double dz, double diskR, zd eps,
                   double tx, double ty, const double* surface, double lambda) {
    hipStream_t s = ctx->stream;
    Grid g;
    g.upload(coords, s);
    const size_t K = g.K();
    DevBuf dE(sizeof(zd) * K), dS, dScr, dfd, dph;
    upload(dE.p, E, sizeof(zd) * K, s);
    if (surface && kind == BF_PROP_PROPAGATOR) {
        dS.alloc(sizeof(zd) * K);
        upload(dS.p, surface, sizeof(zd) * K, s);
    }
    PropEngine<double> pe(ctx, &g, kind, diskR, dS.as<zd>());
    RegionParams rp{eps, dz, tx, ty, lambda, dS.p ? 0 : -1};
    pe.build_tables({rp});
    FieldDesc fd{0, 0, pe.rtab_of.empty() ? -1 : pe.rtab_of[0], 0, 0};
    dfd.alloc(sizeof(FieldDesc));
    upload(dfd.p, &fd, sizeof(fd), s);
    zd ph = phase_1d(eps, lambda, dz);
    dph.alloc(sizeof(zd));
    upload(dph.p, &ph, sizeof(zd), s);
    if (kind == BF_PROP_MOMENTUM) dScr.alloc(sizeof(zd) * K);
    pe.propagate(dE.as<zd>(), dfd.as<FieldDesc>(), 1, dph.as<zd>(), dScr.as<zd>(), nullptr);
    download(E, dE.p, sizeof(zd) * K, s);
    BF_HIP(hipStreamSynchronize(s));
}

// dance_intro amplitudes (src/boost_fractor_recurprop.jl:171-198), host scalar setup.
// eps: the regions dance_intro sees; returns ax[n*2 + slot].
static zd cdiv_(zd a, zd b) {
    if (fabs(b.re) >= fabs(b.im)) {
        double r = b.im / b.re, d = b.re + b.im * r;
        return cmk<double>((a.re + a.im * r) / d, (a.im - a.re * r) / d);
    }
    double r = b.re / b.im, d = b.re * r + b.im;
    return cmk<double>((a.re * r + a.im) / d, (a.im * r - a.re) / d);
}

std::vector<zd> dance_intro_amplitudes(int nd, const zd* eps, zd r1) {
    std::vector<zd> ax(2 * (size_t)nd);
    const zd one = cmk<double>(1.0, 0.0);
    for (int n = 0; n < nd; n++) {
        zd axr;
        if (n == 0) {
            axr = (r1.re == 1.0 && r1.im == 0.0) ? cmk<double>(-1.0, 0.0) : cmk<double>(0.0, 0.0);
        } else {
            zd ei = eps[n - 1], em = eps[n];
            zd den = cadd(cmul(ei, csqrt_j(em)), cmul(em, csqrt_j(ei)));
            axr = cdiv_(cmul(csqrt_j(em), csub(one, cdiv_(ei, em))), den);
        }
        zd ei = eps[n];
        zd em = (n == nd - 1) ? one : eps[n + 1];
        zd den = cadd(cmul(ei, csqrt_j(em)), cmul(em, csqrt_j(ei)));
        zd axl = cdiv_(cmul(csqrt_j(ei), csub(one, cdiv_(em, ei))), den);
        ax[2 * n] = axr;
        ax[2 * n + 1] = axl;
    }
    return ax;
}

void dance_intro_dev(bf_ctx* ctx, const bf_coords* coords, int nd, const double* eps, zd r1, double diskR,
                     double* out) {
    hipStream_t s = ctx->stream;
    Grid g;
    g.upload(coords, s);
    const size_t K = g.K();
    std::vector<zd> ax = dance_intro_amplitudes(nd, reinterpret_cast<const zd*>(eps), r1);
    DevBuf dax(sizeof(zd) * ax.size()), dout(sizeof(zd) * K * 2 * nd);
    upload(dax.p, ax.data(), sizeof(zd) * ax.size(), s);
    hipLaunchKernelGGL(k_dance_intro<double>, dim3(grid1d(K, 256, 512), 2 * nd, 1), dim3(256), 0, s,
                       dout.as<zd>(), nd, K, g.nx, dax.as<zd>(), g.X.as<double>(), g.Y.as<double>(), diskR * diskR,
                       (size_t)0);
    BF_CHECK_LAUNCH();
    download(out, dout.p, sizeof(zd) * K * 2 * nd, s);
    BF_HIP(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------------------
// dancer plan
// ---------------------------------------------------------------------------
template <typename T>
struct DancerPlan : DancerPlanBase {
    Grid g;
    int kind, R, nmax, nfreq, fb, K_;
    double amin, diskR;
    std::vector<double> freqs;
    std::vector<zd> eps;  // trimmed: regions 1..nreg-1 of the setup
    std::vector<double> dist, tx, ty;
    bool has_surf = false;
    DevBuf surf, coef, fields, scratch, Eout, hist, power, active, amax, count, fdesc, ridx, ax, emit, reflect, ph1d;
    std::unique_ptr<PropEngine<T>> pe;
    DancerOptions opt;
    int nslices;
    std::vector<FieldDesc> fd_host;
    // timing: HIP events recorded around every launch of the row and column
    // passes on the plan's stream, reduced once per frequency batch
    bool timing = false;
    std::vector<hipEvent_t> ev;  // [bounce][kernel][start/stop]
    double ms[2] = {0, 0};
    int64_t nl[2] = {0, 0};

    ~DancerPlan() override {
        for (auto e : ev) (void)hipEventDestroy(e);
    }

    void tmark(int n, int k, int which) {
        if (timing) BF_HIP(hipEventRecord(ev[((size_t)n * 2 + k) * 2 + which], ctx->stream));
    }
    void tcollect(int nb_bounces) {
        if (!timing || nb_bounces <= 0) return;
        BF_HIP(hipEventSynchronize(ev[((size_t)(nb_bounces - 1) * 2 + 0) * 2 + 1]));
        for (int n = 0; n < nb_bounces; n++)
            for (int k = 0; k < 2; k++) {
                float m = 0.f;
                BF_HIP(hipEventElapsedTime(&m, ev[((size_t)n * 2 + k) * 2], ev[((size_t)n * 2 + k) * 2 + 1]));
                ms[k] += m;
                nl[k] += 1;
            }
    }
    void set_timing(bool on) override {
        timing = on;
        if (on && ev.empty() && nmax >= 0) {
            ev.resize((size_t)(nmax + 1) * 4);
            for (auto& e : ev) BF_HIP(hipEventCreate(&e));
        }
        ms[0] = ms[1] = 0;
        nl[0] = nl[1] = 0;
    }
    void kernel_times(double* m, int64_t* l) override {
        for (int k = 0; k < 2; k++) {
            if (m) m[k] = ms[k];
            if (l) l[k] = nl[k];
        }
    }

    void init(bf_ctx* c, const bf_setup* st, const bf_coords* co, double amin_, int nmax_, const double* fr,
              int nfr, int kind_, double diskR_, const DancerOptions& o) {
        ctx = c;
        opt = o;
        kind = kind_;
        amin = amin_;
        nmax = nmax_;
        diskR = diskR_;
        nfreq = nfr;
        freqs.assign(fr, fr + nfr);
        hipStream_t s = ctx->stream;
        g.upload(co, s);
        const size_t K = g.K();
        R = st->nreg - 1;
        // trimmed copies (src/boost_fractor_recurprop.jl:28-34)
        const zd* e = reinterpret_cast<const zd*>(st->eps);
        eps.assign(e + 1, e + st->nreg);
        dist.assign(st->distance + 1, st->distance + st->nreg);
        tx.assign(st->tilt_x + 1, st->tilt_x + st->nreg);
        ty.assign(st->tilt_y + 1, st->tilt_y + st->nreg);
        if (st->surfaces && kind == BF_PROP_PROPAGATOR) {
            has_surf = true;
            surf.alloc(sizeof(zd) * K * R);
            upload(surf.p, st->surfaces + 2 * K, sizeof(zd) * K * R, s);
        }
        std::vector<zd> r(R + 1);
        const zd* rr = reinterpret_cast<const zd*>(st->r);
        for (int i = 0; i < R; i++) r[i] = rr[i];
        r[R] = cmk<double>(0.0, 0.0);  // append!(bdry.r, 0.0)
        coef.alloc(sizeof(zd) * (R + 1));
        upload(coef.p, r.data(), sizeof(zd) * (R + 1), s);
        // frequency batch: keep the fields of a batch around 192 MB (Infinity Cache sized)
        const size_t per = sizeof(cplx<T>) * K * 2 * R;
        fb = opt.freq_batch > 0 ? opt.freq_batch : (int)std::max<size_t>(1, (192u << 20) / std::max<size_t>(per, 1));
        fb = std::max(1, std::min(fb, std::min(nfreq, 64)));
        nslices = opt.immediatesum ? 1 : nmax + 1;
        fields.alloc(per * fb);
        if (kind == BF_PROP_MOMENTUM) scratch.alloc(per * fb);
        Eout.alloc(sizeof(zd) * K * nfreq);
        if (nslices > 1) hist.alloc(sizeof(zd) * K * (size_t)nslices * fb);
        power.alloc(sizeof(double) * nfreq);
        active.alloc(sizeof(int) * fb);
        amax.alloc(sizeof(double) * fb);
        count.alloc(sizeof(int) * nfreq);
        BF_HIP(hipMemsetAsync(count.p, 0, sizeof(int) * nfreq, s));
        // field descriptors: field j = (f, region, slot), table index f*R + region
        std::vector<FieldDesc> fd((size_t)fb * 2 * R);
        for (int f = 0; f < fb; f++)
            for (int q = 0; q < 2 * R; q++) {
                FieldDesc& d = fd[(size_t)f * 2 * R + q];
                d.off = (int64_t)((size_t)f * 2 * R + q) * (int64_t)K;
                d.ktab = f * R + q / 2;
                d.rtab = -1;
                d.group = f;
                d.pad = 0;
            }
        fdesc.alloc(sizeof(FieldDesc) * fd.size());
        upload(fdesc.p, fd.data(), sizeof(FieldDesc) * fd.size(), s);
        fd_host = fd;
        ridx.alloc(sizeof(int) * (size_t)fb * R);
        ph1d.alloc(sizeof(zd) * (size_t)fb * 2 * R);
        if (opt.emit) {
            emit.alloc(sizeof(zd) * K * 2 * R);
            upload(emit.p, opt.emit, sizeof(zd) * K * 2 * R, s);
        } else if (opt.reflect) {
            reflect.alloc(sizeof(zd) * K);
            upload(reflect.p, opt.reflect, sizeof(zd) * K, s);
        } else {
            std::vector<zd> a = dance_intro_amplitudes(R, eps.data(), r[0]);
            ax.alloc(sizeof(zd) * a.size());
            upload(ax.p, a.data(), sizeof(zd) * a.size(), s);
        }
        pe.reset(new PropEngine<T>(ctx, &g, kind, diskR, surf.as<zd>()));
        pe->col_tx = opt.col_tx;
        BF_HIP(hipStreamSynchronize(s));
    }

    void run() override {
        hipStream_t s = ctx->stream;
        const size_t K = g.K();
        const bool fused = (kind == BF_PROP_PROPAGATOR || kind == BF_PROP_NOTILTS);
        for (int f0 = 0; f0 < nfreq; f0 += fb) {
            const int nb = std::min(fb, nfreq - f0);
            // tables for this batch: table f*R + region
            std::vector<RegionParams> regs;
            regs.reserve((size_t)fb * R);
            for (int f = 0; f < fb; f++) {
                const double lam = 299792458.0 / freqs[std::min(f0 + f, nfreq - 1)];
                for (int i = 0; i < R; i++)
                    regs.push_back(RegionParams{eps[i], dist[i], tx[i], ty[i], lam, has_surf ? (int64_t)(K * i) : -1});
            }
            pe->build_tables(regs);
            std::vector<int> ri((size_t)fb * R);
            for (size_t q = 0; q < ri.size(); q++) ri[q] = pe->rtab_of.empty() ? -1 : pe->rtab_of[q];
            upload(ridx.p, ri.data(), sizeof(int) * ri.size(), s);
            if (!fused) {  // the unfused propagators read the tilt/surface table index from the descriptors
                for (auto& d : fd_host) d.rtab = ri[d.ktab];
                upload(fdesc.p, fd_host.data(), sizeof(FieldDesc) * fd_host.size(), s);
            }
            if (kind == BF_PROP_1D) {
                std::vector<zd> ph((size_t)fb * 2 * R);
                for (int f = 0; f < fb; f++)
                    for (int q = 0; q < 2 * R; q++)
                        ph[(size_t)f * 2 * R + q] = phase_1d(eps[q / 2], regs[(size_t)f * R].lambda, dist[q / 2]);
                upload(ph1d.p, ph.data(), sizeof(zd) * ph.size(), s);
            }
            // initial fields
            const size_t fstride = K * 2 * R;
            if (opt.emit) {
                hipLaunchKernelGGL(k_broadcast<T>, dim3(grid1d(fstride, 256, 2048), nb), dim3(256), 0, s,
                                   fields.as<cplx<T>>(), emit.as<zd>(), fstride, fstride);
            } else if (opt.reflect) {
                BF_HIP(hipMemsetAsync(fields.p, 0, sizeof(cplx<T>) * fstride * nb, s));
                // fields[R-1][leftmoving = slot 1] = reflect
                hipLaunchKernelGGL(k_broadcast<T>, dim3(grid1d(K, 256, 1024), nb), dim3(256), 0, s,
                                   fields.as<cplx<T>>() + ((size_t)(R - 1) * 2 + 1) * K, reflect.as<zd>(), K, fstride);
            } else {
                hipLaunchKernelGGL(k_dance_intro<T>, dim3(grid1d(K, 256, 512), 2 * R, nb), dim3(256), 0, s,
                                   fields.as<cplx<T>>(), R, K, g.nx, ax.as<zd>(), g.X.as<double>(), g.Y.as<double>(),
                                   diskR * diskR, fstride);
            }
            BF_CHECK_LAUNCH();
            std::vector<int> act(fb, 0);
            for (int f = 0; f < nb; f++) act[f] = (1.0 >= amin) ? 1 : 0;
            upload(active.p, act.data(), sizeof(int) * fb, s);
            BF_HIP(hipMemsetAsync(amax.p, 0, sizeof(double) * fb, s));
            zd* eout = Eout.as<zd>() + (size_t)f0 * K;
            zd* eo_target = (nslices > 1) ? hist.as<zd>() : eout;
            if (nslices == 1) BF_HIP(hipMemsetAsync(eout, 0, sizeof(zd) * K * nb, s));
            else BF_HIP(hipMemsetAsync(hist.p, 0, sizeof(zd) * K * nslices * nb, s));
            const int nf = nb * 2 * R;
            const FieldDesc* fd = fdesc.as<FieldDesc>();
            int rm = 0;
            if (fused) {
                pe->rows(fields.as<cplx<T>>(), fd, nf, RF_MASK | RF_FWD, active.as<int>());
                for (int n = 0; n <= nmax; n++) {
                    tmark(n, 1, 0);
                    pe->cols(fields.as<cplx<T>>(), fields.as<cplx<T>>(), fd, nf, CF_FWD | CF_KMID | CF_INV,
                             active.as<int>());
                    tmark(n, 1, 1);
                    DancerRowArgs<T> a;
                    a.fields = fields.as<cplx<T>>();
                    a.R = R; a.nx = g.nx; a.ny = g.ny; a.rm = rm; a.tw = pe->twx;
                    a.rtab = pe->has_rtab() ? pe->rtab.template as<cplx<T>>() : nullptr;
                    a.rtab_idx = ridx.as<int>();
                    a.coef = coef.as<zd>();
                    a.X = g.X.as<double>(); a.Y = g.Y.as<double>(); a.R2 = diskR * diskR;
                    a.flags = (n < nmax ? DF_FWD : 0u) | (amin != 0.0 ? DF_AMAX : 0u) |
                              (nslices > 1 ? DF_EOUT_ASSIGN : 0u);
                    a.Eout = eo_target;
                    a.eout_fstride = K * nslices;
                    a.slice = nslices > 1 ? n : 0;
                    a.amax = amax.as<double>();
                    a.active = active.as<int>();
                    a.count = count.as<int>() + f0;
                    tmark(n, 0, 0);
                    dispatch_n<DancerRowLaunch, T>(g.nx, a, nb, opt.rows, s);
                    tmark(n, 0, 1);
                    if (amin != 0.0) {
                        hipLaunchKernelGGL(k_amin_check, dim3(1), dim3(64), 0, s, active.as<int>(), amax.as<double>(),
                                           amin, nb);
                        BF_CHECK_LAUNCH();
                    }
                    rm = 1 - rm;
                }
                tcollect(nmax + 1);
            } else {
                for (int n = 0; n <= nmax; n++) {
                    pe->propagate(fields.as<cplx<T>>(), fd, nf, ph1d.as<zd>(), scratch.as<cplx<T>>(),
                                  active.as<int>());
                    hipLaunchKernelGGL(k_dancer_update<T>, dim3(grid1d(K, 256, 256), R + 1, nb), dim3(256), 0, s,
                                       fields.as<cplx<T>>(), R, K, rm, coef.as<zd>(), eo_target, K * nslices,
                                       nslices > 1 ? n : 0, nslices > 1 ? 1 : 0,
                                       amin != 0.0 ? amax.as<double>() : nullptr, active.as<int>(),
                                       count.as<int>() + f0);
                    BF_CHECK_LAUNCH();
                    if (amin != 0.0) {
                        hipLaunchKernelGGL(k_amin_check, dim3(1), dim3(64), 0, s, active.as<int>(), amax.as<double>(),
                                           amin, nb);
                        BF_CHECK_LAUNCH();
                    }
                    rm = 1 - rm;
                }
            }
            if (nslices > 1) {
                if (opt.returnsum) {
                    hipLaunchKernelGGL(k_reverse_sum, dim3(grid1d(K, 256, 512), nb), dim3(256), 0, s, eout,
                                       hist.as<zd>(), nslices, K);
                    BF_CHECK_LAUNCH();
                } else {
                    // history is returned as is: stage it in the caller-visible order later
                    BF_HIP(hipMemcpyAsync(hist_out_ptr(f0), hist.p, sizeof(zd) * K * nslices * nb,
                                          hipMemcpyDeviceToDevice, s));
                }
            }
            hipLaunchKernelGGL(k_power, dim3(nb), dim3(1024), 0, s, eout, K, power.as<double>() + f0);
            BF_CHECK_LAUNCH();
        }
    }

    // full history storage for returnsum == 0 with nmax+1 slices
    DevBuf hist_all;
    zd* hist_out_ptr(int f0) {
        const size_t K = g.K();
        if (!hist_all.p) hist_all.alloc(sizeof(zd) * K * nslices * nfreq);
        return hist_all.as<zd>() + (size_t)f0 * nslices * K;
    }

    void fetch(double* Eo, double* pw, int32_t* bounces) override {
        hipStream_t s = ctx->stream;
        const size_t K = g.K();
        if (Eo) {
            if (!opt.returnsum && nslices > 1)
                download(Eo, hist_all.p, sizeof(zd) * K * nslices * nfreq, s);
            else
                download(Eo, Eout.p, sizeof(zd) * K * nfreq, s);
        }
        if (pw) download(pw, power.p, sizeof(double) * nfreq, s);
        if (bounces) download(bounces, count.p, sizeof(int) * nfreq, s);
        BF_HIP(hipStreamSynchronize(s));
    }
};

DancerPlanBase* make_dancer_plan(bf_ctx* ctx, const bf_setup* st, const bf_coords* co, double amin, int nmax,
                                 const double* fr, int nfr, int kind, double diskR, int precision,
                                 const DancerOptions& o) {
    if (precision == BF_PREC_MIXED) {
        auto* p = new DancerPlan<float>();
        try { p->init(ctx, st, co, amin, nmax, fr, nfr, kind, diskR, o); } catch (...) { delete p; throw; }
        return p;
    }
    auto* p = new DancerPlan<double>();
    try { p->init(ctx, st, co, amin, nmax, fr, nfr, kind, diskR, o); } catch (...) { delete p; throw; }
    return p;
}

// ---------------------------------------------------------------------------
// cheerleader (src/boost_fractor_recurprop.jl:226-363), Gauss-Seidel sweep,
// batched over frequencies.  fields[f][Nd][K].
// ---------------------------------------------------------------------------
// F[i+1] += t*F[i]; F[i] *= r      (prologue and last region, :285-286, :343-344)
template <typename T>
__global__ void k_ch_push(cplx<T>* F, int Nd, size_t K, int i, zd t_, zd r_, const int* active) {
    const int f = blockIdx.y;
    if (active && !active[f]) return;
    const cplx<T> t = cmk<T>((T)t_.re, (T)t_.im), r = cmk<T>((T)r_.re, (T)r_.im);
    cplx<T>* Fi = F + ((size_t)f * Nd + i) * K;
    cplx<T>* Fn = Fi + K;
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < K; p += (size_t)gridDim.x * blockDim.x) {
        const cplx<T> v = Fi[p];
        Fn[p] = cadd(Fn[p], cmul(t, v));
        Fi[p] = cmul(v, r);
    }
}
// F[0] += tl*F[1]; F[1] *= rl    (first region, :318-320)
template <typename T>
__global__ void k_ch_first(cplx<T>* F, int Nd, size_t K, zd tl_, zd rl_, const int* active) {
    const int f = blockIdx.y;
    if (active && !active[f]) return;
    const cplx<T> tl = cmk<T>((T)tl_.re, (T)tl_.im), rl = cmk<T>((T)rl_.re, (T)rl_.im);
    cplx<T>* F0 = F + (size_t)f * Nd * K;
    cplx<T>* F1 = F0 + K;
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < K; p += (size_t)gridDim.x * blockDim.x) {
        const cplx<T> v = F1[p];
        F0[p] = cadd(F0[p], cmul(tl, v));
        F1[p] = cmul(v, rl);
    }
}
// pair update at boundary i (:330-334)
template <typename T>
__global__ void k_ch_pair(cplx<T>* F, int Nd, size_t K, int i, zd rl_, zd tr_, zd rr_, zd tl_, const int* active) {
    const int f = blockIdx.y;
    if (active && !active[f]) return;
    const cplx<T> rl = cmk<T>((T)rl_.re, (T)rl_.im), tr = cmk<T>((T)tr_.re, (T)tr_.im);
    const cplx<T> rr = cmk<T>((T)rr_.re, (T)rr_.im), tl = cmk<T>((T)tl_.re, (T)tl_.im);
    cplx<T>* Fi = F + ((size_t)f * Nd + i) * K;
    cplx<T>* Fn = Fi + K;
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < K; p += (size_t)gridDim.x * blockDim.x) {
        const cplx<T> a = Fi[p], c = Fn[p];
        Fn[p] = cadd(cmul(c, rl), cmul(tr, a));
        Fi[p] = cadd(cmul(a, rr), cmul(tl, c));
    }
}
// emission of the left-moving fields (:290-300)
template <typename T>
__global__ void k_ch_emit_left(cplx<T>* F, const cplx<T>* __restrict__ E, int Nd, size_t K, zd tl_, zd rl_) {
    const int f = blockIdx.y;
    const cplx<T> tl = cmk<T>((T)tl_.re, (T)tl_.im), rl = cmk<T>((T)rl_.re, (T)rl_.im);
    cplx<T>* Fb = F + (size_t)f * Nd * K;
    const cplx<T>* Eb = E + (size_t)f * Nd * 2 * K;
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < K; p += (size_t)gridDim.x * blockDim.x) {
        for (int i = 0; i <= Nd - 2; i++) Fb[(size_t)i * K + p] = cadd(Fb[(size_t)i * K + p], Eb[((size_t)i * 2 + 1) * K + p]);
        const cplx<T> e = Eb[((size_t)(Nd - 1) * 2 + 1) * K + p];
        Fb[(size_t)(Nd - 2) * K + p] = cadd(Fb[(size_t)(Nd - 2) * K + p], cmul(tl, e));
        Fb[(size_t)(Nd - 1) * K + p] = cadd(Fb[(size_t)(Nd - 1) * K + p], cmul(rl, e));
    }
}
template <typename T>
__global__ void k_ch_copy_right(cplx<T>* F, const cplx<T>* __restrict__ E, int Nd, size_t K) {
    const int f = blockIdx.y;
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < (size_t)Nd * K; p += (size_t)gridDim.x * blockDim.x) {
        const size_t i = p / K, q = p % K;
        F[(size_t)f * Nd * K + p] = E[((size_t)f * Nd + i) * 2 * K + q];
    }
}
template <typename T>
__global__ void k_amax_fields(const cplx<T>* F, size_t n, double* amax, const int* active) {
    const int f = blockIdx.y;
    if (active && !active[f]) return;
    double m = 0.0;
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
        const cplx<T> v = F[(size_t)f * n + p];
        m = fmax(m, hypot((double)v.re, (double)v.im));
    }
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0 && m > 0.0) atomic_max_pos(amax + f, m);
}

template <typename T>
void cheerleader_run(bf_ctx* ctx, const bf_setup* st, const bf_coords* co, double amin, int nmax, const double* fr,
                     int nfr, int kind, const double* emit_h, const double* reflect_h, double diskR, double* out_right,
                     double* out_left) {
    hipStream_t s = ctx->stream;
    Grid g;
    g.upload(co, s);
    const size_t K = g.K();
    const int Nd = st->nreg;
    DevBuf surf;
    if (st->surfaces && kind == BF_PROP_PROPAGATOR) {
        surf.alloc(sizeof(zd) * K * Nd);
        upload(surf.p, st->surfaces, sizeof(zd) * K * Nd, s);
    }
    const zd* eps = reinterpret_cast<const zd*>(st->eps);
    const zd* r = reinterpret_cast<const zd*>(st->r);
    std::vector<zd> rl(Nd), rr(Nd), tl(Nd), tr(Nd);
    for (int i = 0; i < Nd - 1; i++) {
        rl[i] = cmk<double>(-r[i].re, -r[i].im);
        rr[i] = r[i];
        tl[i] = cmk<double>(1.0 - r[i].re, -r[i].im);
        tr[i] = cmk<double>(1.0 + r[i].re, r[i].im);
    }
    const size_t per = sizeof(cplx<T>) * K * Nd;
    int fb = (int)std::max<size_t>(1, (192u << 20) / std::max<size_t>(per * 3, 1));
    fb = std::max(1, std::min(fb, std::min(nfr, 64)));
    DevBuf F(per * fb), EM(per * 2 * fb), scr, dax, demit, dref, out(sizeof(zd) * K * fb), act(sizeof(int) * fb),
        amx(sizeof(double) * fb), dfd, dph;
    if (kind == BF_PROP_MOMENTUM) scr.alloc(per * fb);
    PropEngine<T> pe(ctx, &g, kind, diskR, surf.as<zd>());
    // descriptors: per region i, fields (f, i) for all f; pairs (f, i), (f, i+1)
    std::vector<FieldDesc> fd;
    std::vector<int> single_off(Nd), pair_off(Nd);
    for (int i = 0; i < Nd; i++) {
        single_off[i] = (int)fd.size();
        for (int f = 0; f < fb; f++) fd.push_back(FieldDesc{(int64_t)(((size_t)f * Nd + i) * K), f * Nd + i, -1, f, 0});
    }
    for (int i = 0; i + 1 < Nd; i++) {
        pair_off[i] = (int)fd.size();
        for (int f = 0; f < fb; f++)
            for (int q = 0; q < 2; q++)
                fd.push_back(FieldDesc{(int64_t)(((size_t)f * Nd + i + q) * K), f * Nd + i + q, -1, f, 0});
    }
    dfd.alloc(sizeof(FieldDesc) * fd.size());
    dph.alloc(sizeof(zd) * fd.size());
    if (emit_h) {
        demit.alloc(sizeof(zd) * K * 2 * Nd);
        upload(demit.p, emit_h, sizeof(zd) * K * 2 * Nd, s);
    } else if (reflect_h) {
        dref.alloc(sizeof(zd) * K);
        upload(dref.p, reflect_h, sizeof(zd) * K, s);
    } else {
        // dance_intro on the copy with distance[2:end], eps[2:end] and the original r (:262-267)
        std::vector<zd> a = dance_intro_amplitudes(Nd - 1, eps + 1, r[0]);
        dax.alloc(sizeof(zd) * a.size());
        upload(dax.p, a.data(), sizeof(zd) * a.size(), s);
    }
    const unsigned gb = grid1d(K, 256, 1024);
    for (int f0 = 0; f0 < nfr; f0 += fb) {
        const int nb = std::min(fb, nfr - f0);
        std::vector<RegionParams> regs;
        for (int f = 0; f < fb; f++) {
            const double lam = 299792458.0 / fr[std::min(f0 + f, nfr - 1)];
            for (int i = 0; i < Nd; i++)
                regs.push_back(RegionParams{eps[i], st->distance[i], st->tilt_x[i], st->tilt_y[i], lam,
                                            surf.p ? (int64_t)(K * i) : -1});
        }
        pe.build_tables(regs);
        std::vector<zd> ph(fd.size());
        for (size_t q = 0; q < fd.size(); q++) {
            fd[q].rtab = pe.rtab_of.empty() ? -1 : pe.rtab_of[fd[q].ktab];
            const RegionParams& rp = regs[fd[q].ktab];
            ph[q] = phase_1d(rp.eps, rp.lambda, rp.dz);
        }
        upload(dfd.p, fd.data(), sizeof(FieldDesc) * fd.size(), s);
        upload(dph.p, ph.data(), sizeof(zd) * ph.size(), s);
        // emit[f][Nd][2][K]
        const size_t estride = K * 2 * Nd;
        if (emit_h) {
            hipLaunchKernelGGL(k_broadcast<T>, dim3(grid1d(estride, 256, 2048), nb), dim3(256), 0, s,
                               EM.as<cplx<T>>(), demit.as<zd>(), estride, estride);
        } else {
            BF_HIP(hipMemsetAsync(EM.p, 0, per * 2 * nb, s));
            if (reflect_h)
                hipLaunchKernelGGL(k_broadcast<T>, dim3(gb, nb), dim3(256), 0, s,
                                   EM.as<cplx<T>>() + ((size_t)(Nd - 1) * 2 + 1) * K, dref.as<zd>(), K, estride);
            else
                hipLaunchKernelGGL(k_dance_intro<T>, dim3(grid1d(K, 256, 512), 2 * (Nd - 1), nb), dim3(256), 0, s,
                                   EM.as<cplx<T>>() + 2 * K, Nd - 1, K, g.nx, dax.as<zd>(), g.X.as<double>(),
                                   g.Y.as<double>(), diskR * diskR, estride);
        }
        BF_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_ch_copy_right<T>, dim3(grid1d(K * Nd, 256, 4096), nb), dim3(256), 0, s, F.as<cplx<T>>(),
                           EM.as<cplx<T>>(), Nd, K);
        BF_CHECK_LAUNCH();
        std::vector<int> a0(fb, 0);
        for (int f = 0; f < nb; f++) a0[f] = (1.0 >= amin) ? 1 : 0;
        const FieldDesc* D = dfd.as<FieldDesc>();
        const zd* P1 = dph.as<zd>();
        cplx<T>* Fp = F.as<cplx<T>>();
        cplx<T>* Sp = scr.as<cplx<T>>();
        auto prop_single = [&](int i, const int* active) {
            pe.propagate(Fp, D + single_off[i], nb, P1 + single_off[i], Sp, active);
        };
        auto prop_pair = [&](int i, const int* active) {
            pe.propagate(Fp, D + pair_off[i], 2 * nb, P1 + pair_off[i], Sp, active);
        };
        // prologue (:281-300), runs for every frequency slot
        for (int i = 1; i <= Nd - 2; i++) {
            prop_single(i, nullptr);
            hipLaunchKernelGGL(k_ch_push<T>, dim3(gb, nb), dim3(256), 0, s, Fp, Nd, K, i, tr[i], rr[i], nullptr);
            BF_CHECK_LAUNCH();
        }
        hipLaunchKernelGGL(k_ch_emit_left<T>, dim3(gb, nb), dim3(256), 0, s, Fp, EM.as<cplx<T>>(), Nd, K, tl[Nd - 2],
                           rl[Nd - 2]);
        BF_CHECK_LAUNCH();
        upload(act.p, a0.data(), sizeof(int) * fb, s);
        BF_HIP(hipMemsetAsync(amx.p, 0, sizeof(double) * fb, s));
        const int* A = act.as<int>();
        for (int n = 0; n <= nmax; n++) {
            prop_single(1, A);
            hipLaunchKernelGGL(k_ch_first<T>, dim3(gb, nb), dim3(256), 0, s, Fp, Nd, K, tl[0], rl[0], A);
            BF_CHECK_LAUNCH();
            for (int i = 1; i <= Nd - 3; i++) {
                prop_pair(i, A);
                hipLaunchKernelGGL(k_ch_pair<T>, dim3(gb, nb), dim3(256), 0, s, Fp, Nd, K, i, rl[i], tr[i], rr[i],
                                   tl[i], A);
                BF_CHECK_LAUNCH();
            }
            prop_single(Nd - 2, A);
            hipLaunchKernelGGL(k_ch_push<T>, dim3(gb, nb), dim3(256), 0, s, Fp, Nd, K, Nd - 2, tr[Nd - 2], rr[Nd - 2],
                               A);
            BF_CHECK_LAUNCH();
            if (amin != 0.0) {
                hipLaunchKernelGGL(k_amax_fields<T>, dim3(grid1d(K * Nd, 256, 1024), nb), dim3(256), 0, s, Fp,
                                   K * Nd, amx.as<double>(), A);
                hipLaunchKernelGGL(k_amin_check, dim3(1), dim3(64), 0, s, act.as<int>(), amx.as<double>(), amin, nb);
                BF_CHECK_LAUNCH();
            }
        }
        // outputs: right = F[Nd-1], left = F[0]
        for (int f = 0; f < nb; f++) {
            hipLaunchKernelGGL(k_to_double<T>, dim3(gb), dim3(256), 0, s, out.as<zd>() + (size_t)f * K,
                               Fp + ((size_t)f * Nd + Nd - 1) * K, K);
        }
        BF_CHECK_LAUNCH();
        download(out_right + 2 * K * f0, out.p, sizeof(zd) * K * nb, s);
        if (out_left) {
            BF_HIP(hipStreamSynchronize(s));
            for (int f = 0; f < nb; f++)
                hipLaunchKernelGGL(k_to_double<T>, dim3(gb), dim3(256), 0, s, out.as<zd>() + (size_t)f * K,
                                   Fp + ((size_t)f * Nd) * K, K);
            BF_CHECK_LAUNCH();
            download(out_left + 2 * K * f0, out.p, sizeof(zd) * K * nb, s);
        }
        BF_HIP(hipStreamSynchronize(s));
    }
}

// ---------------------------------------------------------------------------
// helpers for the mode-matching unit
// ---------------------------------------------------------------------------
// unnormalised 2D FFT of nf contiguous fields (optionally aperture-masked first)
void fft_fields(bf_ctx* ctx, const Grid& g, zd* base, int nf, bool mask, double diskR) {
    PropEngine<double> pe(ctx, &g, BF_PROP_NOTILTS, diskR, nullptr);
    std::vector<FieldDesc> fd(nf);
    for (int j = 0; j < nf; j++) fd[j] = FieldDesc{(int64_t)j * (int64_t)g.K(), -1, -1, 0, 0};
    DevBuf dfd(sizeof(FieldDesc) * nf);
    upload(dfd.p, fd.data(), sizeof(FieldDesc) * nf, ctx->stream);
    pe.rows(base, dfd.as<FieldDesc>(), nf, (mask ? RF_MASK : 0u) | RF_FWD, nullptr);
    pe.cols(base, base, dfd.as<FieldDesc>(), nf, CF_FWD, nullptr);
    BF_HIP(hipStreamSynchronize(ctx->stream));
}

// prop(...) applied to nf contiguous fields with one region's parameters
void propagate_fields(bf_ctx* ctx, const Grid& g, int kind, double diskR, const zd* surfaces, zd* base, int nf,
                      zd eps, double dz, double tx, double ty, int64_t surf_off, double lambda, zd* scratch) {
    PropEngine<double> pe(ctx, &g, kind, diskR, surfaces);
    pe.build_tables({RegionParams{eps, dz, tx, ty, lambda, surf_off}});
    const int rt = pe.rtab_of.empty() ? -1 : pe.rtab_of[0];
    std::vector<FieldDesc> fd(nf);
    for (int j = 0; j < nf; j++) fd[j] = FieldDesc{(int64_t)j * (int64_t)g.K(), 0, rt, 0, 0};
    std::vector<zd> ph(nf, phase_1d(eps, lambda, dz));
    DevBuf dfd(sizeof(FieldDesc) * nf), dph(sizeof(zd) * nf);
    upload(dfd.p, fd.data(), sizeof(FieldDesc) * nf, ctx->stream);
    upload(dph.p, ph.data(), sizeof(zd) * nf, ctx->stream);
    pe.propagate(base, dfd.as<FieldDesc>(), nf, dph.as<zd>(), scratch, nullptr);
    BF_HIP(hipStreamSynchronize(ctx->stream));
}

void build_ktabs(bf_ctx* ctx, const Grid& g, const std::vector<KTabDesc>& d, double scale, zd* out) {
    if (d.empty()) return;
    DevBuf dd(sizeof(KTabDesc) * d.size());
    upload(dd.p, d.data(), sizeof(KTabDesc) * d.size(), ctx->stream);
    hipLaunchKernelGGL(k_build_ktab<double>, dim3(grid1d(g.K(), 256, 1024), (unsigned)d.size()), dim3(256), 0,
                       ctx->stream, out, dd.as<KTabDesc>(), g.nx, g.ny, g.kX.as<double>(), g.kY.as<double>(), scale);
    BF_CHECK_LAUNCH();
    BF_HIP(hipStreamSynchronize(ctx->stream));
}

void build_rtabs(bf_ctx* ctx, const Grid& g, const std::vector<RTabDesc>& d, const zd* surfaces, zd* out) {
    if (d.empty()) return;
    DevBuf dd(sizeof(RTabDesc) * d.size());
    upload(dd.p, d.data(), sizeof(RTabDesc) * d.size(), ctx->stream);
    hipLaunchKernelGGL(k_build_rtab<double>, dim3(grid1d(g.K(), 256, 1024), (unsigned)d.size()), dim3(256), 0,
                       ctx->stream, out, dd.as<RTabDesc>(), g.nx, g.ny, g.X.as<double>(), g.Y.as<double>(), surfaces);
    BF_CHECK_LAUNCH();
    BF_HIP(hipStreamSynchronize(ctx->stream));
}

// ---------------------------------------------------------------------------
// stream bandwidth probe
// ---------------------------------------------------------------------------
__global__ void k_stream_copy(double2* __restrict__ dst, const double2* __restrict__ src, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

double stream_copy_probe(bf_ctx* ctx, int64_t bytes, int reps) {
    hipStream_t s = ctx->stream;
    const size_t n = (size_t)bytes / sizeof(double2);
    DevBuf a(n * sizeof(double2)), b(n * sizeof(double2));
    BF_HIP(hipMemsetAsync(a.p, 0, n * sizeof(double2), s));
    hipEvent_t e0, e1;
    BF_HIP(hipEventCreate(&e0));
    BF_HIP(hipEventCreate(&e1));
    const unsigned blocks = 256 * 16;
    hipLaunchKernelGGL(k_stream_copy, dim3(blocks), dim3(256), 0, s, b.as<double2>(), a.as<double2>(), n);
    BF_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; r++)
        hipLaunchKernelGGL(k_stream_copy, dim3(blocks), dim3(256), 0, s, b.as<double2>(), a.as<double2>(), n);
    BF_CHECK_LAUNCH();
    BF_HIP(hipEventRecord(e1, s));
    BF_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    BF_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return (double)ms / std::max(reps, 1);
}

template void cheerleader_run<double>(bf_ctx*, const bf_setup*, const bf_coords*, double, int, const double*, int, int,
                                      const double*, const double*, double, double*, double*);
template void cheerleader_run<float>(bf_ctx*, const bf_setup*, const bf_coords*, double, int, const double*, int, int,
                                     const double*, const double*, double, double*, double*);

}  // namespace bf
