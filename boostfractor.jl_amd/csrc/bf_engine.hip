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
#include "bf_fftreg.h"

#include <algorithm>
#include <chrono>
#include <array>
#include <type_traits>
#include <cmath>
#include <cstring>

namespace bf {

// ---------------------------------------------------------------------------
// descriptors
// ---------------------------------------------------------------------------
enum : uint32_t { RF_INV = 1, RF_RTAB = 2, RF_MASK = 4, RF_FWD = 8 };
enum : uint32_t { CF_FWD = 1, CF_KMID = 2, CF_INV = 4, CF_KLOAD = 8, CF_INSHIFT = 16, CF_OUTSHIFT = 32 };
enum : uint32_t { DF_FWD = 1, DF_AMAX = 2, DF_EOUT_ASSIGN = 4 };
enum : int { FD_WHOLE = 1 };  // FieldDesc.pad: column pass stores every row of this field
constexpr int FD_PAIR_SHIFT = 8;  // FieldDesc.pad >> FD_PAIR_SHIFT: region pair of a cheerleader pair step

// Precision tags: `double` = FP64 storage and arithmetic; `float` = fields stored
// in FP32 in HBM, all in-register arithmetic (FFTs, phases, boundary updates) in
// FP64; `f32pure` = FP32 storage and FP32 arithmetic.
template <typename T> struct Storage { using type = T; };
template <> struct Storage<f32pure> { using type = float; };
template <typename T> struct Compute { using type = double; };
template <> struct Compute<f32pure> { using type = float; };
#define BF_TYPES(T)                           \
    using S = typename Storage<T>::type;      \
    using C = typename Compute<T>::type

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

// quarter != 0 (unshifted tables of even-sized grids whose k range is mirror
// symmetric, k[j] = -k[N-1-j] as -dk(N-1)/2 : dk : dk(N-1)/2 is): only bins
// mx < nx/2, my < ny/2 are stored, as [ny/2][nx/2]; bin nx-1-mx reads entry mx
// (and likewise in y), whose k differs from its own by the rounding of the range.
template <typename T>
__global__ void k_build_ktab(cplx<T>* out, const KTabDesc* __restrict__ d, int nx, int ny,
                             const double* __restrict__ kX, const double* __restrict__ kY, double scale,
                             int quarter = 0) {
    const int tab = blockIdx.y;
    const int wx = quarter ? nx / 2 : nx;
    const size_t K = (size_t)wx * (quarter ? ny / 2 : ny);
    const KTabDesc dd = d[tab];
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < K; i += (size_t)gridDim.x * blockDim.x) {
        int mx = (int)(i % wx), my = (int)(i / wx);
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
__global__ void k_scale_fields(typename Storage<T>::type* base_, const FieldDesc* __restrict__ fd,
                               const zd* __restrict__ ph, size_t K, const int* __restrict__ active) {
    BF_TYPES(T);
    cplx<S>* base = reinterpret_cast<cplx<S>*>(base_);
    const int j = blockIdx.y;
    const FieldDesc d = fd[j];
    if (active && !active[d.group]) return;
    const cplx<C> p = cmk<C>((C)ph[j].re, (C)ph[j].im);
    cplx<S>* f = base + d.off;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < K; i += (size_t)gridDim.x * blockDim.x)
        f[i] = ccast<S>(cmul(ccast<C>(f[i]), p));
}

// ---------------------------------------------------------------------------
// row pass
// ---------------------------------------------------------------------------
template <typename T>
struct RowArgs {
    cplx<typename Storage<T>::type>* base;
    const FieldDesc* fd;
    int nf, nx, ny;
    const cplx<typename Compute<T>::type>* tw;
    uint32_t flags;
    const cplx<typename Compute<T>::type>* rtab;
    const double* X;
    const double* Y;
    double R2;
    const int* active;
    // rows outside [ylo, yhi) are skipped (not read, not written) when yhi > ylo
    int ylo, yhi;
};

template <typename T, int N>
__global__ void __launch_bounds__(1024) k_rows(RowArgs<T> a) {
    BF_TYPES(T);
    extern __shared__ unsigned char smem_raw[];
    cplx<C>* smem = reinterpret_cast<cplx<C>*>(smem_raw);
    constexpr int TPL = LineCfg<N>::TPL;
    const int n = (N >= 16) ? N : a.nx;
    const int lpb = blockDim.x / TPL;
    const int li = threadIdx.x / TPL, t = threadIdx.x % TPL;
    cplx<C>* s = smem + (size_t)li * n;
    cplx<C>* s2 = smem + (size_t)(lpb + li) * n;
    const long nlines = (long)a.nf * a.ny;
    const long line = (long)blockIdx.x * lpb + li;
    bool valid = line < nlines;
    const int j = valid ? (int)(line / a.ny) : 0;
    const int y = valid ? (int)(line % a.ny) : 0;
    const FieldDesc d = a.fd[j];
    if (valid && a.active && !a.active[d.group]) valid = false;
    if (a.yhi > a.ylo && (y < a.ylo || y >= a.yhi)) valid = false;
    cplx<S>* row = a.base + d.off + (size_t)y * a.nx;
    for (int x = t; x < n; x += TPL) s[x] = valid ? ccast<C>(row[x]) : cmk<C>(0, 0);
    __syncthreads();
    if (a.flags & RF_INV) line_fft<C, N, true>(s, s2, n, t, a.tw);
    if (a.flags & (RF_RTAB | RF_MASK)) {
        const cplx<C>* tab = ((a.flags & RF_RTAB) && d.rtab >= 0)
                                 ? a.rtab + ((size_t)d.rtab * a.ny + y) * a.nx
                                 : nullptr;
        const double yy = a.Y[y];
        for (int x = t; x < n; x += TPL) {
            cplx<C> v = s[x];
            if (tab) v = cmul(v, tab[x]);
            if (a.flags & RF_MASK) {
                const double xx = a.X[x];
                if (!(fabs(xx * xx + yy * yy) < a.R2)) v = cmk<C>(0, 0);
            }
            s[x] = v;
        }
        __syncthreads();
    }
    if (a.flags & RF_FWD) line_fft<C, N, false>(s, s2, n, t, a.tw);
    if (valid)
        for (int x = t; x < n; x += TPL) row[x] = ccast<S>(s[x]);
}

// ---------------------------------------------------------------------------
// column pass
// ---------------------------------------------------------------------------
template <typename T>
struct ColArgs {
    const cplx<typename Storage<T>::type>* src;
    cplx<typename Storage<T>::type>* dst;
    const FieldDesc* fd;
    int nf, nx, ny;
    const cplx<typename Compute<T>::type>* tw;
    uint32_t flags;
    const cplx<typename Compute<T>::type>* ktab;
    int tx;  // columns per block
    const int* active;
    // Row band: rows outside [ylo, yhi) are known to be zero on input (they are
    // not loaded) and are not needed on output (not stored), except for the
    // fields whose descriptor carries FD_WHOLE, stored whole.
    int ylo, yhi;
    int kq;  // 0: full k-space tables; else quarter tables of row width kq = nx/2 (dancer propagation)
    int all_whole;  // every field is stored whole (the band limits the loads only)
};

template <typename T>
__device__ __forceinline__ bool in_band(const ColArgs<T>& a, int y) {
    return (unsigned)(y - a.ylo) < (unsigned)(a.yhi - a.ylo);
}
__device__ __forceinline__ bool store_whole(const FieldDesc& d) { return (d.pad & FD_WHOLE) != 0; }

template <typename T, int N>
__global__ void __launch_bounds__(1024) k_cols(ColArgs<T> a) {
    BF_TYPES(T);
    extern __shared__ unsigned char smem_raw[];
    cplx<C>* smem = reinterpret_cast<cplx<C>*>(smem_raw);
    constexpr int TPL = LineCfg<N>::TPL;
    const int n = (N >= 16) ? N : a.ny;
    const int ld = n + 1;
    const int tx = a.tx;
    cplx<C>* s = smem;
    cplx<C>* s2 = smem + (size_t)tx * ld;
    const FieldDesc d = a.fd[blockIdx.y];
    if (a.active && !a.active[d.group]) return;  // uniform over the block
    const int x0 = blockIdx.x * tx;
    const int nx = a.nx, ny = a.ny;
    const int hx = nx / 2, hy = ny / 2;
    const cplx<S>* src = a.src + d.off;
    cplx<S>* dst = a.dst + d.off;
    const cplx<C>* kt = (d.ktab >= 0) ? a.ktab + (size_t)d.ktab * ny * nx : nullptr;
    const uint32_t fl = a.flags;
    const bool whole = store_whole(d) || a.all_whole;
    for (int i = threadIdx.x; i < tx * n; i += blockDim.x) {
        const int c = i % tx, my = i / tx, mx = x0 + c;
        cplx<C> v = cmk<C>(0, 0);
        if (mx < nx) {
            int sx = mx, sy = my;
            if (fl & CF_INSHIFT) {  // ifftshift on load
                sx = (mx + hx) % nx;
                sy = (my + hy) % ny;
            }
            if (in_band(a, sy)) {
                const size_t o = (size_t)sy * nx + sx;
                v = ccast<C>(src[o]);
                if ((fl & CF_KLOAD) && kt) v = cmul(v, kt[o]);
            }
        }
        s[c * ld + my] = v;
    }
    __syncthreads();
    const int c = threadIdx.x / TPL, t = threadIdx.x % TPL;
    cplx<C>* line = s + c * ld;
    cplx<C>* line2 = s2 + c * ld;
    if (fl & CF_FWD) line_fft<C, N, false>(line, line2, n, t, a.tw);
    if ((fl & CF_KMID) && kt) {
        for (int i = threadIdx.x; i < tx * n; i += blockDim.x) {
            const int cc = i % tx, my = i / tx, mx = x0 + cc;
            if (mx < nx) s[cc * ld + my] = cmul(s[cc * ld + my], kt[(size_t)my * nx + mx]);
        }
        __syncthreads();
    }
    if (fl & CF_INV) line_fft<C, N, true>(line, line2, n, t, a.tw);
    for (int i = threadIdx.x; i < tx * n; i += blockDim.x) {
        const int cc = i % tx, my = i / tx, mx = x0 + cc;
        if (mx < nx) {
            int dx = mx, dy = my;
            if (fl & CF_OUTSHIFT) {  // fftshift on store
                dx = (mx + hx) % nx;
                dy = (my + hy) % ny;
            }
            if (whole || in_band(a, dy)) dst[(size_t)dy * nx + dx] = ccast<S>(s[cc * ld + my]);
        }
    }
}

// ---------------------------------------------------------------------------
// dancer: fused row pass with the boundary update
// ---------------------------------------------------------------------------
template <typename T>
struct DancerRowArgs {
    cplx<typename Storage<T>::type>* fields;  // [fb][R][2][ny][nx]
    int R, nx, ny;
    int rm;  // slot holding the right-moving fields this bounce
    const cplx<typename Compute<T>::type>* tw;
    const cplx<typename Compute<T>::type>* rtab;
    const int* rtab_idx;  // [fb][R], -1: none
    const zd* coef;       // boundary reflectivities r_b, b = 0..R (r_R = 0 appended)
    const double* X;
    const double* Y;
    double R2;
    const uint32_t* keep_tab;  // [ny][nx/16] aperture bits of the register path, or null
    int keep_band;             // band class of keep_tab: no bit outside [keep_band, 15 - keep_band]
    uint32_t flags;
    zd* Eout;  // [nslices][ny][nx] per frequency slot, stride eout_fstride
    size_t eout_fstride;
    int slice;
    double* amax;  // [fb]
    const int* active;
    int* count;  // [fb] bounce counters
    // Row band [ylo, yhi): outside it the next bounce's aperture zeroes every
    // row, so those rows are only needed for Eout at the exit boundary.  The
    // launcher fills the tile map from the band and its rows per block.
    int ylo, yhi;
    int tile0, nband, ntiles;
};

// Block -> (boundary b, row tile).  Blocks [0, nband (R+1)) cover the row tiles
// that meet the band at every boundary; the remaining ntiles - nband blocks
// cover the other row tiles at the exit boundary only (exit_only: Eout is
// accumulated, nothing is stored because the aperture zeroes those rows).
template <typename T>
__device__ __forceinline__ void dancer_block(const DancerRowArgs<T>& a, int i, int& b, int& tile, bool& exit_only) {
    const int nb = a.nband * (a.R + 1);
    if (i < nb) {
        b = i / a.nband;
        tile = a.tile0 + i % a.nband;
        exit_only = false;
    } else {
        const int j = i - nb;
        b = a.R;
        tile = j < a.tile0 ? j : j + a.nband;
        exit_only = true;
    }
}

__device__ __forceinline__ void atomic_max_pos(double* addr, double v) {
    // non-negative doubles order like their bit patterns
    atomicMax(reinterpret_cast<unsigned long long*>(addr), (unsigned long long)__double_as_longlong(v));
}

template <typename T, int N>
__global__ void __launch_bounds__(1024) k_dancer_rows(DancerRowArgs<T> a) {
    BF_TYPES(T);
    extern __shared__ unsigned char smem_raw[];
    cplx<C>* smem = reinterpret_cast<cplx<C>*>(smem_raw);
    __shared__ double red[32];
    constexpr int TPL = LineCfg<N>::TPL;
    const int n = (N >= 16) ? N : a.nx;
    const int f = blockIdx.y;
    int b, tile;
    bool exit_only;
    dancer_block(a, blockIdx.x, b, tile, exit_only);
    if (a.active && !a.active[f]) return;  // uniform over the block
    if (a.count && blockIdx.x == 0 && threadIdx.x == 0) a.count[f] += 1;
    const int lpb = blockDim.x / TPL;  // 2 * rows per block
    const int li = threadIdx.x / TPL, t = threadIdx.x % TPL;
    const int side = li & 1;  // 0: A = region b-1 right-moving, 1: B = region b left-moving
    const int y = tile * (lpb >> 1) + (li >> 1);
    const int reg = side == 0 ? b - 1 : b;
    const int slot = side == 0 ? a.rm : 1 - a.rm;
    const bool exists = reg >= 0 && reg < a.R && y < a.ny;
    const size_t K = (size_t)a.nx * a.ny;
    cplx<C>* s = smem + (size_t)li * n;
    cplx<C>* s2 = smem + (size_t)(lpb + li) * n;
    cplx<S>* row = a.fields + (((size_t)f * a.R + (exists ? reg : 0)) * 2 + slot) * K + (size_t)(exists ? y : 0) * a.nx;
    for (int x = t; x < n; x += TPL) s[x] = exists ? ccast<C>(row[x]) : cmk<C>(0, 0);
    __syncthreads();
    line_fft<C, N, true>(s, s2, n, t, a.tw);
    if (a.rtab) {
        const int ti = exists ? a.rtab_idx[f * a.R + reg] : -1;
        if (ti >= 0) {
            const cplx<C>* tab = a.rtab + (size_t)ti * K + (size_t)y * a.nx;
            for (int x = t; x < n; x += TPL) s[x] = cmul(s[x], tab[x]);
        }
        __syncthreads();
    }
    // boundary update (src/boost_fractor_recurprop.jl:90-119), result staged in s2
    const zd r = a.coef[b];
    const cplx<C> rT = cmk<C>((C)r.re, (C)r.im);
    const cplx<C> mrT = cmk<C>((C)(-r.re), (C)(-r.im));
    const cplx<C> omrT = cmk<C>((C)(1.0 - r.re), (C)(-r.im));
    const cplx<C> oprT = cmk<C>((C)(1.0 + r.re), (C)r.im);
    const cplx<C>* p = smem + (size_t)(li ^ 1) * n;  // partner line, same y
    double vmax = 0.0;
    const double yy = a.Y[exists ? y : 0];
    for (int x = t; x < n; x += TPL) {
        const cplx<C> v = s[x], w = p[x];
        cplx<C> nv = cmk<C>(0, 0);
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
                if (!(fabs(xx * xx + yy * yy) < a.R2)) nv = cmk<C>(0, 0);  // next bounce's aperture
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
    if (!(a.flags & DF_FWD) || exit_only) return;  // only Eout is needed
    for (int x = t; x < n; x += TPL) s[x] = s2[x];
    __syncthreads();
    line_fft<C, N, false>(s, s2, n, t, a.tw);
    if (exists)
        for (int x = t; x < n; x += TPL) row[x] = ccast<S>(s[x]);
}

// ---------------------------------------------------------------------------
// register-FFT versions of the passes (power-of-two lines)
// ---------------------------------------------------------------------------
template <typename T, int N>
struct RegE {
    // 16 points per thread: a line of <= 256 points needs a single LDS exchange
    static constexpr int E = 16;
    static constexpr int TT = N / E;
    static constexpr int LD = N + N / E;
    // Column-pass line stride.  The threads of a wave span 16 columns, so the
    // stride decides how the columns' lines fall on the LDS banks; LD itself
    // (a multiple of the bank row) puts all 16 on the same banks.  Measured on
    // gfx950 (SQ_LDS_BANK_CONFLICT per active LDS cycle, 256-point columns):
    // LD 9.5, LD+1 0.56, LD+2 1.1, LD+3 0.56, LD+4 3.8, LD+8 7.3 -> odd stride.
    static constexpr int LDC = LD + 1;
};

// Aperture bits of the E points x = t + TT m of row y: bit m set when the
// point lies inside (all coordinate loads are issued together).
template <int E, int TT>
__device__ __forceinline__ uint32_t aperture_bits(const double* __restrict__ X, double yy, double R2, int t) {
    double xs[E];
#pragma unroll
    for (int m = 0; m < E; m++) xs[m] = X[t + TT * m];
    uint32_t bits = 0;
#pragma unroll
    for (int m = 0; m < E; m++) bits |= (in_aperture(xs[m], yy, R2) ? 1u : 0u) << m;
    return bits;
}

template <typename T, int N>
__global__ void __launch_bounds__(512) k_rows_reg(RowArgs<T> a) {
    BF_TYPES(T);
    extern __shared__ unsigned char smem_raw[];
    constexpr int E = RegE<T, N>::E, TT = RegE<T, N>::TT, LD = RegE<T, N>::LD;
    cplx<S>* smem = reinterpret_cast<cplx<S>*>(smem_raw);
    const int lpb = blockDim.x / TT;
    const int li = threadIdx.x / TT, t = threadIdx.x % TT;
    cplx<S>* s = smem + (size_t)li * LD;
    // the grid covers the rows [ylo, yhi) of every field (all rows when yhi <= ylo)
    const bool band = a.yhi > a.ylo;
    const int nr = band ? a.yhi - a.ylo : a.ny, y0 = band ? a.ylo : 0;
    const long nlines = (long)a.nf * nr;
    const long line = (long)blockIdx.x * lpb + li;
    bool valid = line < nlines;
    const int j = valid ? (int)(line / nr) : 0;
    const int y = valid ? y0 + (int)(line % nr) : 0;
    const FieldDesc d = a.fd[j];
    if (valid && a.active && !a.active[d.group]) valid = false;
    const uint32_t K = (uint32_t)N * (uint32_t)a.ny;
    const auto rf = buf_rsrc(a.base + d.off, K * (uint32_t)sizeof(cplx<S>));
    const uint32_t o0 = (uint32_t)y * N + t;
    cplx<C> v[E];
    {
        cplx<S> raw[E];
#pragma unroll
        for (int m = 0; m < E; m++) raw[m] = buf_load<S>(rf, valid ? (o0 + TT * m) * (uint32_t)sizeof(cplx<S>) : BUF_OOB);
#pragma unroll
        for (int m = 0; m < E; m++) v[m] = ccast<C>(raw[m]);
    }
    if (a.flags & RF_INV) fft_reg<C, S, N, E, true>(v, s, t, a.tw);
    if ((a.flags & RF_RTAB) && d.rtab >= 0) {
        const auto rt = buf_rsrc(a.rtab + (size_t)d.rtab * K, K * (uint32_t)sizeof(cplx<C>));
        cplx<C> tab[E];
#pragma unroll
        for (int m = 0; m < E; m++) tab[m] = buf_load<C>(rt, (o0 + TT * m) * (uint32_t)sizeof(cplx<C>));
#pragma unroll
        for (int m = 0; m < E; m++) v[m] = cmul(v[m], tab[m]);
    }
    if (a.flags & RF_MASK) {
        const uint32_t keep = aperture_bits<E, TT>(a.X, a.Y[y], a.R2, t);
#pragma unroll
        for (int m = 0; m < E; m++)
            if (!((keep >> m) & 1u)) v[m] = cmk<C>(0, 0);
    }
    if (a.flags & RF_FWD) fft_reg<C, S, N, E, false>(v, s, t, a.tw);
#pragma unroll
    for (int m = 0; m < E; m++)
        buf_store<S>(rf, valid ? (o0 + TT * m) * (uint32_t)sizeof(cplx<S>) : BUF_OOB, ccast<S>(v[m]));
}

// FL != 0: the flags are fixed at compile time (the dancer's FWD|KMID|INV).
// BM > 0 (dancer propagation only): every row of the band [ylo, yhi) is t + TT m
// with BM <= m <= 15 - BM, so the forward transform skips the butterflies of the
// other (zero) elements and the inverse transform forms the band elements
// first; a field stored whole then forms the others from the same last-pass
// input, still in LDS.
template <typename T, int N, uint32_t FL, int BM = 0>
__global__ void __launch_bounds__(512) k_cols_reg(ColArgs<T> a) {
    BF_TYPES(T);
    extern __shared__ unsigned char smem_raw[];
    constexpr int E = RegE<T, N>::E, TT = RegE<T, N>::TT, LDC = RegE<T, N>::LDC;
    cplx<S>* smem = reinterpret_cast<cplx<S>*>(smem_raw);
    const int jf = blockIdx.y;
    const FieldDesc d = a.fd[jf];
    if (a.active && !a.active[d.group]) return;  // uniform over the block
    const uint32_t fl = FL ? FL : a.flags;
    const int tx = a.tx;
    const int c = threadIdx.x % tx, t = threadIdx.x / tx;
    const int nx = a.nx;
    const int mx = blockIdx.x * tx + c;
    const bool ok = mx < nx;
    const uint32_t K = (uint32_t)N * (uint32_t)nx;
    const auto rs = buf_rsrc(a.src + d.off, K * (uint32_t)sizeof(cplx<S>));
    const auto rd = buf_rsrc(a.dst + d.off, K * (uint32_t)sizeof(cplx<S>));
    const bool has_kt = d.ktab >= 0;
    const uint32_t KT = a.kq ? (uint32_t)a.kq * (uint32_t)(N / 2) : K;  // elements per table
    const auto rk = buf_rsrc(a.ktab + (size_t)(has_kt ? d.ktab : 0) * KT, KT * (uint32_t)sizeof(cplx<C>));
    const bool whole = store_whole(d) || a.all_whole;
    const int sx = (fl & CF_INSHIFT) ? (mx + nx / 2) % nx : mx;
    cplx<S>* s = smem + (size_t)c * LDC;
    cplx<C> v[E];
    constexpr uint32_t PROPF = CF_FWD | CF_KMID | CF_INV;
    if constexpr (FL == PROPF) {
        // The dancer's propagation (no shifts): element m of this thread is row
        // y = t + TT m.  A resource over the band rows only answers rows outside
        // [ylo, yhi) with zero: their offsets, (y - ylo) * nx + mx elements,
        // wrap to >= 2^31 below the band and reach past its end above it, so
        // every offset is the lane base plus a uniform multiple of 16 rows.
        const uint32_t rowb = (uint32_t)nx * (uint32_t)sizeof(cplx<S>), stepb = (uint32_t)TT * rowb;
        const uint32_t bandb = (uint32_t)(a.yhi - a.ylo) * rowb;
        const auto rsb = buf_rsrc(a.src + d.off + (size_t)a.ylo * nx, bandb);
        const uint32_t lband = ok ? (uint32_t)((t - a.ylo) * nx + mx) * (uint32_t)sizeof(cplx<S>) : BUF_OOB;
        constexpr int MLO = BM, MHI = E - 1 - BM;
        cplx<S> raw[E];
#pragma unroll
        for (int m = MLO; m <= MHI; m++) raw[m] = buf_load<S>(rsb, lband + (uint32_t)m * stepb);
#pragma unroll
        for (int m = 0; m < E; m++) v[m] = (m >= MLO && m <= MHI) ? ccast<C>(raw[m]) : cmk<C>(0, 0);
        if constexpr (BM == 0)
            fft_reg<C, S, N, E, false>(v, s, t, a.tw);
        else
            fft_reg_m<C, S, N, E, false, band_mask16(BM), 0xFFFF>(v, s, t, a.tw);
        if (has_kt) {
            constexpr int KCH = sizeof(C) == 8 ? 8 : E;
            if (a.kq) {
                // quarter table: column min(mx, nx-1-mx); row y = t + TT m is < N/2 for
                // m < E/2 (entry y) and >= N/2 otherwise (entry N-1-y = TT-1-t + TT (15-m))
                const uint32_t krow = (uint32_t)a.kq * (uint32_t)sizeof(cplx<C>), kstep = (uint32_t)TT * krow;
                const int qx = mx < nx / 2 ? mx : nx - 1 - mx;
                const uint32_t lo = ok ? (uint32_t)(t * a.kq + qx) * (uint32_t)sizeof(cplx<C>) : BUF_OOB;
                const uint32_t hi = ok ? (uint32_t)((TT - 1 - t) * a.kq + qx) * (uint32_t)sizeof(cplx<C>) : BUF_OOB;
#pragma unroll
                for (int m0 = 0; m0 < E; m0 += KCH) {
                    cplx<C> k[KCH];
#pragma unroll
                    for (int q = 0; q < KCH; q++) {
                        const int m = m0 + q;
                        k[q] = buf_load<C>(rk, m < E / 2 ? lo + (uint32_t)m * kstep : hi + (uint32_t)(E - 1 - m) * kstep);
                    }
#pragma unroll
                    for (int q = 0; q < KCH; q++) v[m0 + q] = cmul(v[m0 + q], k[q]);
                }
            } else {
                const uint32_t krow = (uint32_t)nx * (uint32_t)sizeof(cplx<C>), kstep = (uint32_t)TT * krow;
                const uint32_t lk = ok ? (uint32_t)(t * nx + mx) * (uint32_t)sizeof(cplx<C>) : BUF_OOB;
#pragma unroll
                for (int m0 = 0; m0 < E; m0 += KCH) {
                    cplx<C> k[KCH];
#pragma unroll
                    for (int q = 0; q < KCH; q++) k[q] = buf_load<C>(rk, lk + (uint32_t)(m0 + q) * kstep);
#pragma unroll
                    for (int q = 0; q < KCH; q++) v[m0 + q] = cmul(v[m0 + q], k[q]);
                }
            }
        }
        if constexpr (BM == 0)
            fft_reg<C, S, N, E, true>(v, s, t, a.tw);
        else
            fft_reg_m<C, S, N, E, true, 0xFFFF, band_mask16(BM)>(v, s, t, a.tw);
        // band rows, or every row of a field stored whole
        const auto rdb = buf_rsrc(a.dst + d.off + (whole ? 0 : (size_t)a.ylo * nx), whole ? K * (uint32_t)sizeof(cplx<S>) : bandb);
        const uint32_t lst = ok ? (uint32_t)((t - (whole ? 0 : a.ylo)) * nx + mx) * (uint32_t)sizeof(cplx<S>) : BUF_OOB;
#pragma unroll
        for (int m = MLO; m <= MHI; m++) buf_store<S>(rdb, lst + (uint32_t)m * stepb, ccast<S>(v[m]));
        if constexpr (BM > 0) {
            if (whole) {  // uniform over the block
                fft_reg_last_m<C, S, N, E, true, 0xFFFF & ~band_mask16(BM)>(v, s, t, a.tw);
#pragma unroll
                for (int m = 0; m < E; m++)
                    if (m < MLO || m > MHI) buf_store<S>(rdb, lst + (uint32_t)m * stepb, ccast<S>(v[m]));
            }
        }
        return;
    }
    {
        uint32_t off[E];
#pragma unroll
        for (int m = 0; m < E; m++) {
            const int my = t + TT * m;
            const int sy = (fl & CF_INSHIFT) ? ((my + N / 2) & (N - 1)) : my;
            off[m] = (ok && in_band(a, sy)) ? (uint32_t)(sy * nx + sx) : BUF_OOB;
        }
        cplx<S> raw[E];
#pragma unroll
        for (int m = 0; m < E; m++)
            raw[m] = buf_load<S>(rs, off[m] == BUF_OOB ? BUF_OOB : off[m] * (uint32_t)sizeof(cplx<S>));
        if ((fl & CF_KLOAD) && has_kt) {
            cplx<C> k[E];
#pragma unroll
            for (int m = 0; m < E; m++)
                k[m] = buf_load<C>(rk, off[m] == BUF_OOB ? BUF_OOB : off[m] * (uint32_t)sizeof(cplx<C>));
#pragma unroll
            for (int m = 0; m < E; m++) v[m] = cmul(ccast<C>(raw[m]), k[m]);
        } else {
#pragma unroll
            for (int m = 0; m < E; m++) v[m] = ccast<C>(raw[m]);
        }
    }
    if (fl & CF_FWD) fft_reg<C, S, N, E, false>(v, s, t, a.tw);
    if ((fl & CF_KMID) && has_kt) {
        // in chunks of KCH loads: bounds the registers held by loads in flight
        constexpr int KCH = sizeof(C) == 8 ? 8 : E;
#pragma unroll
        for (int m0 = 0; m0 < E; m0 += KCH) {
            cplx<C> k[KCH];
#pragma unroll
            for (int q = 0; q < KCH; q++)
                k[q] = buf_load<C>(rk, ok ? (uint32_t)((t + TT * (m0 + q)) * nx + mx) * (uint32_t)sizeof(cplx<C>)
                                          : BUF_OOB);
#pragma unroll
            for (int q = 0; q < KCH; q++) v[m0 + q] = cmul(v[m0 + q], k[q]);
        }
    }
    if (fl & CF_INV) fft_reg<C, S, N, E, true>(v, s, t, a.tw);
    const int dx = (fl & CF_OUTSHIFT) ? (mx + nx / 2) % nx : mx;
#pragma unroll
    for (int m = 0; m < E; m++) {
        int dy = t + TT * m;
        if (fl & CF_OUTSHIFT) dy = (dy + N / 2) & (N - 1);
        const bool st = ok && (whole || in_band(a, dy));
        buf_store<S>(rd, st ? (uint32_t)(dy * nx + dx) * (uint32_t)sizeof(cplx<S>) : BUF_OOB, ccast<S>(v[m]));
    }
}

// OBM > 0: only elements m in [OBM, 15 - OBM] of the inverse transform are
// formed, updated and exchanged (blocks whose update needs no other element:
// not the exit boundary, no field maximum).  IBM > 0: the aperture keeps no
// element outside [IBM, 15 - IBM], so the forward transform skips the others.
template <typename T, int N, int OBM, int IBM>
__device__ __forceinline__ void dancer_rows_reg_body(const DancerRowArgs<T>& a) {
    BF_TYPES(T);
    extern __shared__ unsigned char smem_raw[];
    constexpr int E = RegE<T, N>::E, TT = RegE<T, N>::TT, LD = RegE<T, N>::LD;
    cplx<S>* smem = reinterpret_cast<cplx<S>*>(smem_raw);
    __shared__ double red[32];
    const int f = blockIdx.y;
    int b, tile;
    bool exit_only;
    dancer_block(a, blockIdx.x, b, tile, exit_only);
    if (a.active && !a.active[f]) return;  // uniform over the block
    if (a.count && blockIdx.x == 0 && threadIdx.x == 0) a.count[f] += 1;
    const int lpb = blockDim.x / TT;  // 2 * rows per block
    const int li = threadIdx.x / TT, t = threadIdx.x % TT;
    const int side = li & 1;  // 0: A = region b-1 right-moving, 1: B = region b left-moving
    const int y = tile * (lpb >> 1) + (li >> 1);
    const int reg = side == 0 ? b - 1 : b;
    const int slot = side == 0 ? a.rm : 1 - a.rm;
    const bool exists = reg >= 0 && reg < a.R && y < a.ny;
    const int yc = y < a.ny ? y : a.ny - 1;
    const uint32_t K = (uint32_t)N * (uint32_t)a.ny;
    cplx<S>* s = smem + (size_t)li * LD;
    cplx<S>* sp = smem + (size_t)(li ^ 1) * LD;  // partner line, same y
    // A wave spans several lines (different regions and slots), so the line
    // address is per lane: plain global accesses under one per-line condition
    // (a buffer resource per line would not be wave-uniform).
    cplx<S>* line = a.fields + (((size_t)f * a.R + (exists ? reg : 0)) * 2 + slot) * K + (size_t)yc * N + t;
    const uint32_t o0 = (uint32_t)yc * N + t;
    cplx<C> v[E];
    {
        cplx<S> raw[E];
        if (exists) {
#pragma unroll
            for (int m = 0; m < E; m++) raw[m] = line[TT * m];
        } else {
#pragma unroll
            for (int m = 0; m < E; m++) raw[m] = cmk<S>(0, 0);
        }
#pragma unroll
        for (int m = 0; m < E; m++) v[m] = ccast<C>(raw[m]);
    }
    // next bounce's aperture (bit m: keep point t + TT m)
    const uint32_t keep = (a.flags & DF_FWD) ? (a.keep_tab ? a.keep_tab[yc * TT + t] : aperture_bits<E, TT>(a.X, a.Y[yc], a.R2, t))
                                             : ~0u;
    // elements formed by the inverse transform
    constexpr int OLO = OBM, OHI = E - 1 - OBM;
    if constexpr (OBM == 0)
        fft_reg<C, S, N, E, true>(v, s, t, a.tw);
    else
        fft_reg_m<C, S, N, E, true, 0xFFFF, band_mask16(OBM)>(v, s, t, a.tw);
    if (a.rtab) {
        const int ti = exists ? a.rtab_idx[f * a.R + reg] : -1;
        if (ti >= 0) {
            const cplx<C>* tl = a.rtab + (size_t)ti * K + o0;
            cplx<C> tab[E];
#pragma unroll
            for (int m = OLO; m <= OHI; m++) tab[m] = tl[TT * m];
#pragma unroll
            for (int m = OLO; m <= OHI; m++) v[m] = cmul(v[m], tab[m]);
        }
    }
    // exchange with the partner line through LDS (the FFT's buffers are free now)
    __syncthreads();
#pragma unroll
    for (int m = OLO; m <= OHI; m++) s[t + TT * m] = ccast<S>(v[m]);
    __syncthreads();
    // boundary update (src/boost_fractor_recurprop.jl:90-119) as own*c_own + partner*c_w:
    //   side 0 (A):  b < R: A' = r A + (1-r) B;    b = R: A' = r A, Eout += (1+r) A
    //   side 1 (B):  b > 0: B' = -r B + (1+r) A;   b = 0: B' = -r B
    // c_w = 0 exactly where the partner line does not exist (its values are zero).
    const zd r = a.coef[b];
    cplx<C> cown, cw;
    if (side == 0) {
        cown = cmk<C>((C)r.re, (C)r.im);
        cw = (b == a.R) ? cmk<C>(0, 0) : cmk<C>((C)(1.0 - r.re), (C)(-r.im));
    } else {
        cown = cmk<C>((C)(-r.re), (C)(-r.im));
        cw = (b == 0) ? cmk<C>(0, 0) : cmk<C>((C)(1.0 + r.re), (C)r.im);
    }
    if (OBM == 0 && b == a.R) {  // exit boundary (uniform over the block): Eout += (1+r) * A on the A lines
        const bool eo_line = exists && side == 0;
        const bool assign = (a.flags & DF_EOUT_ASSIGN) != 0;
        const zd opr = cmk<double>(1.0 + r.re, r.im);
        if (eo_line) {
            zd* eo = a.Eout + (size_t)f * a.eout_fstride + (size_t)a.slice * K + o0;
            constexpr int ECH = 4;  // exit blocks are few: keep their extra registers low
#pragma unroll
            for (int m0 = 0; m0 < E; m0 += ECH) {
                zd prev[ECH];
#pragma unroll
                for (int q = 0; q < ECH; q++) prev[q] = assign ? cmk<double>(0.0, 0.0) : eo[TT * (m0 + q)];
#pragma unroll
                for (int q = 0; q < ECH; q++)
                    eo[TT * (m0 + q)] = cadd(prev[q], cmul(opr, ccast<double>(v[m0 + q])));
            }
        }
    }
    cplx<S> wp[E];  // partner values, all LDS reads issued together
#pragma unroll
    for (int m = OLO; m <= OHI; m++) wp[m] = sp[t + TT * m];
    double vmax = 0.0;
#pragma unroll
    for (int m = 0; m < E; m++) {
        if (m < OLO || m > OHI) {
            v[m] = cmk<C>(0, 0);  // outside the aperture: zeroed by the mask
            continue;
        }
        const cplx<C> w = ccast<C>(wp[m]);
        cplx<C> nv = exists ? cadd(cmul(v[m], cown), cmul(cw, w)) : cmk<C>(0, 0);
        if (OBM == 0 && (a.flags & DF_AMAX)) vmax = fmax(vmax, hypot((double)nv.re, (double)nv.im));
        if (!((keep >> m) & 1u)) nv = cmk<C>(0, 0);
        v[m] = nv;
    }
    if (OBM == 0 && (a.flags & DF_AMAX)) {
        for (int o = 32; o > 0; o >>= 1) vmax = fmax(vmax, __shfl_xor(vmax, o));
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = vmax;
        __syncthreads();
        if (threadIdx.x == 0) {
            double mm = 0.0;
            for (int w = 0; w < (int)(blockDim.x >> 6); w++) mm = fmax(mm, red[w]);
            if (mm > 0.0) atomic_max_pos(a.amax + f, mm);
        }
    }
    if (!(a.flags & DF_FWD) || exit_only) return;  // only Eout is needed
    if constexpr (IBM == 0)
        fft_reg<C, S, N, E, false>(v, s, t, a.tw);
    else
        fft_reg_m<C, S, N, E, false, band_mask16(IBM), 0xFFFF>(v, s, t, a.tw);
    if (exists)
#pragma unroll
        for (int m = 0; m < E; m++) line[TT * m] = ccast<S>(v[m]);
}

// XB > 0: the aperture bits keep no element outside [XB, 15 - XB] (the launcher
// checks the table and that the forward transform runs)
template <typename T, int N, int XB>
__device__ __forceinline__ void dancer_rows_reg_dispatch(const DancerRowArgs<T>& a) {
    if constexpr (XB == 0) {
        dancer_rows_reg_body<T, N, 0, 0>(a);
    } else {
        int b, tile;
        bool exit_only;
        dancer_block(a, blockIdx.x, b, tile, exit_only);
        if (b == a.R || (a.flags & DF_AMAX))  // uniform over the block
            dancer_rows_reg_body<T, N, 0, XB>(a);
        else
            dancer_rows_reg_body<T, N, XB, XB>(a);
    }
}

template <typename T, int N, int XB = 0>
__global__ void __launch_bounds__(512) k_dancer_rows_reg(DancerRowArgs<T> a) {
    dancer_rows_reg_dispatch<T, N, XB>(a);
}

// The complex64-storage instantiation capped at 128 VGPRs (4 waves per SIMD
// instead of 3, at the price of 4 spilled registers): measured 7% faster; the
// FP64 one would spill far more and is left uncapped.
template <int N, int XB = 0>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) k_dancer_rows_reg4(DancerRowArgs<float> a) {
    dancer_rows_reg_dispatch<float, N, XB>(a);
}

// pointwise bounce update for the unfused propagators (propagator1D, momentum space)
template <typename T>
__global__ void k_dancer_update(cplx<typename Storage<T>::type>* fields, int R, size_t K, int rm,
                                const zd* __restrict__ coef, zd* Eout, size_t eout_fstride, int slice, int assign,
                                double* amax, const int* active, int* count) {
    BF_TYPES(T);
    const int f = blockIdx.z, b = blockIdx.y;
    if (active && !active[f]) return;
    if (count && blockIdx.x == 0 && b == 0 && threadIdx.x == 0) count[f] += 1;
    const zd r = coef[b];
    const cplx<C> rT = cmk<C>((C)r.re, (C)r.im), mrT = cmk<C>((C)(-r.re), (C)(-r.im));
    const cplx<C> omrT = cmk<C>((C)(1.0 - r.re), (C)(-r.im)), oprT = cmk<C>((C)(1.0 + r.re), (C)r.im);
    cplx<S>* A = (b >= 1) ? fields + (((size_t)f * R + (b - 1)) * 2 + rm) * K : nullptr;
    cplx<S>* B = (b < R) ? fields + (((size_t)f * R + b) * 2 + (1 - rm)) * K : nullptr;
    double vmax = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < K; i += (size_t)gridDim.x * blockDim.x) {
        if (b == 0) {
            const cplx<C> nb = cmul(ccast<C>(B[i]), mrT);
            B[i] = ccast<S>(nb);
            vmax = fmax(vmax, hypot((double)nb.re, (double)nb.im));
        } else if (b == R) {
            const cplx<C> va = ccast<C>(A[i]);
            const zd e = cmul(cmk<double>(1.0 + r.re, r.im), ccast<double>(va));
            zd* eo = Eout + (size_t)f * eout_fstride + (size_t)slice * K + i;
            *eo = assign ? e : cadd(*eo, e);
            const cplx<C> na = cmul(va, rT);
            A[i] = ccast<S>(na);
            vmax = fmax(vmax, hypot((double)na.re, (double)na.im));
        } else {
            const cplx<C> va = ccast<C>(A[i]), vb = ccast<C>(B[i]);
            const cplx<C> na = cadd(cmul(va, rT), cmul(omrT, vb));
            const cplx<C> nb = cadd(cmul(vb, mrT), cmul(oprT, va));
            A[i] = ccast<S>(na);
            B[i] = ccast<S>(nb);
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

// one thread per frequency slot of the batch (batches hold up to 256 slots)
static void launch_amin_check(hipStream_t s, int* active, double* amax, double amin, int nb) {
    hipLaunchKernelGGL(k_amin_check, dim3((unsigned)((nb + 63) / 64)), dim3(64), 0, s, active, amax, amin, nb);
}

// dance_intro fill (src/boost_fractor_recurprop.jl:153-201): fields[f][n][s] = ax[n][s] * [x^2+y^2 <= R^2]
// bfield: null -> the default real mask x^2 + y^2 > R^2 ? 0 : 1 (ax .* mask, complex x real);
// else complex [nd+1][K]: region n's right-moving amplitude multiplies layer n, its
// left-moving amplitude layer n+1 (complex x complex, src/boost_fractor_recurprop.jl:196-197).
template <typename T>
__global__ void k_dance_intro(cplx<typename Storage<T>::type>* fields, int nd, size_t K, int nx,
                              const zd* __restrict__ ax, const double* __restrict__ X, const double* __restrict__ Y,
                              double R2, size_t fstride, const zd* __restrict__ bfield) {
#pragma clang fp contract(off)
    BF_TYPES(T);
    (void)sizeof(C);
    const int f = blockIdx.z, q = blockIdx.y;  // q = region*2 + slot
    const zd v = ax[q];
    const zd* layer = bfield ? bfield + (size_t)(q / 2 + (q & 1)) * K : nullptr;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < K; i += (size_t)gridDim.x * blockDim.x) {
        zd e;
        if (layer) {
            e = cmul(v, layer[i]);
        } else {
            const double xx = X[i % nx], yy = Y[i / nx];
            const double bb = (xx * xx + yy * yy > R2) ? 0.0 : 1.0;
            e = cmk<double>(bb * v.re, bb * v.im);
        }
        fields[(size_t)f * fstride + (size_t)q * K + i] = cmk<S>((S)e.re, (S)e.im);
    }
}

// bfield[z][y][x] *= exp(a.re*x + 1im*a.im*x) for every layer: the axion velocity
// phase exp(-1im*Ma*(-velocity_x)*x) with a = (-1im*Ma)*(-velocity_x) (:162-166)
__global__ void k_bfield_velocity(zd* b, size_t n, size_t K, int nx, const double* __restrict__ X, zd a) {
#pragma clang fp contract(off)
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double x = X[(i % K) % nx];
        b[i] = cmul(b[i], cexp_j(cmk<double>(a.re * x, a.im * x)));
    }
}

// broadcast a double field block [nq][K] into every frequency slot of a T buffer
template <typename T>
__global__ void k_broadcast(cplx<typename Storage<T>::type>* dst, const zd* __restrict__ src, size_t n,
                            size_t fstride) {
    BF_TYPES(T);
    (void)sizeof(C);
    const int f = blockIdx.y;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[(size_t)f * fstride + i] = cmk<S>((S)src[i].re, (S)src[i].im);
}

template <typename T>
__global__ void k_to_double(zd* dst, const cplx<typename Storage<T>::type>* __restrict__ src, size_t n) {
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
    hkX.assign(c->kX, c->kX + nx);
    hkY.assign(c->kY, c->kY + ny);
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

static bool use_reg_fft() { return !opts().fft_lds; }

template <typename T, int N>
struct RowLaunch {
    static void run(const RowArgs<T>& a, hipStream_t s) {
        if constexpr (N >= 16) {
            if (use_reg_fft()) {
                using C = RegE<T, N>;
                const int lpb = std::max(1, 256 / C::TT);
                // lines of the row band only (k_rows_reg maps the grid onto [ylo, yhi))
                const long nlines = (long)a.nf * (a.yhi > a.ylo ? a.yhi - a.ylo : a.ny);
                const size_t smem = sizeof(cplx<typename Storage<T>::type>) * (size_t)C::LD * lpb;
                allow_lds((const void*)k_rows_reg<T, N>, smem);
                hipLaunchKernelGGL((k_rows_reg<T, N>), dim3((unsigned)((nlines + lpb - 1) / lpb)), dim3(lpb * C::TT),
                                   smem, s, a);
                BF_CHECK_LAUNCH();
                return;
            }
        }
        const int TPL = LineCfg<N>::TPL;
        const int n = N >= 16 ? N : a.nx;
        const int lpb = std::max(1, 256 / TPL);
        const long nlines = (long)a.nf * a.ny;
        const size_t smem = sizeof(cplx<typename Compute<T>::type>) * (size_t)n * lpb * (N >= 16 ? 1 : 2);
        allow_lds((const void*)k_rows<T, N>, smem);
        hipLaunchKernelGGL((k_rows<T, N>), dim3((unsigned)((nlines + lpb - 1) / lpb)), dim3(lpb * TPL), smem, s, a);
        BF_CHECK_LAUNCH();
    }
};

// Band class of the pruned register transforms: the largest BM in {5, 3} with
// every possibly nonzero element m of a thread in [BM, E-1-BM] (elements mlo..mhi),
// else 0 (no pruning).  Option prune=0 switches the pruned kernels off.
static int prune_class(int mlo, int mhi, int E) {
    if (!opts().prune || mhi < mlo) return 0;
    const int bm = std::min(mlo, E - 1 - mhi);
    return bm >= 5 ? 5 : bm >= 3 ? 3 : 0;
}

template <typename T, int N>
struct ColLaunch {
    static void run(ColArgs<T> a, int tx_req, hipStream_t s) {
        if (a.kq && !(N >= 16 && use_reg_fft() && a.flags == (CF_FWD | CF_KMID | CF_INV) && a.nx == N))
            throw std::runtime_error("quarter k-space tables are read only by the register propagation pass");
        if constexpr (N >= 16) {
            if (use_reg_fft()) {
                using C = RegE<T, N>;
                int tx = tx_req > 0 ? tx_req : (int)(128 / sizeof(cplx<typename Storage<T>::type>));
                tx = std::max(1, std::min(tx, 512 / C::TT));
                while (tx > 1 && sizeof(cplx<typename Storage<T>::type>) * (size_t)tx * C::LDC > 144 * 1024) tx >>= 1;
                a.tx = tx;
                const size_t smem = sizeof(cplx<typename Storage<T>::type>) * (size_t)tx * C::LDC;
                constexpr uint32_t PROP = CF_FWD | CF_KMID | CF_INV;
                if (a.flags == PROP) {
                    const int bm = N >= 128 ? prune_class(a.ylo / C::TT, (a.yhi - 1) / C::TT, C::E) : 0;
                    const dim3 grid((a.nx + tx - 1) / tx, a.nf), block(tx * C::TT);
                    bool pruned = false;
                    if constexpr (N >= 128) {
                        if (bm == 5) {
                            allow_lds((const void*)k_cols_reg<T, N, PROP, 5>, smem);
                            hipLaunchKernelGGL((k_cols_reg<T, N, PROP, 5>), grid, block, smem, s, a);
                            pruned = true;
                        } else if (bm == 3) {
                            allow_lds((const void*)k_cols_reg<T, N, PROP, 3>, smem);
                            hipLaunchKernelGGL((k_cols_reg<T, N, PROP, 3>), grid, block, smem, s, a);
                            pruned = true;
                        }
                    }
                    if (!pruned) {
                        allow_lds((const void*)k_cols_reg<T, N, PROP>, smem);
                        hipLaunchKernelGGL((k_cols_reg<T, N, PROP>), grid, block, smem, s, a);
                    }
                } else {
                    allow_lds((const void*)k_cols_reg<T, N, 0>, smem);
                    hipLaunchKernelGGL((k_cols_reg<T, N, 0>), dim3((a.nx + tx - 1) / tx, a.nf), dim3(tx * C::TT),
                                       smem, s, a);
                }
                BF_CHECK_LAUNCH();
                return;
            }
        }
        const int TPL = LineCfg<N>::TPL;
        const int n = N >= 16 ? N : a.ny;
        int tx = tx_req > 0 ? tx_req : (int)(128 / sizeof(cplx<typename Storage<T>::type>));
        tx = std::max(1, std::min(tx, 1024 / TPL));
        while (tx > 1 && sizeof(cplx<typename Compute<T>::type>) * (size_t)tx * (n + 1) * (N >= 16 ? 1 : 2) > 96 * 1024) tx >>= 1;
        a.tx = tx;
        const size_t smem = sizeof(cplx<typename Compute<T>::type>) * (size_t)tx * (n + 1) * (N >= 16 ? 1 : 2);
        allow_lds((const void*)k_cols<T, N>, smem);
        hipLaunchKernelGGL((k_cols<T, N>), dim3((a.nx + tx - 1) / tx, a.nf), dim3(tx * TPL), smem, s, a);
        BF_CHECK_LAUNCH();
    }
};

template <typename T>
static unsigned dancer_tiles(DancerRowArgs<T>& a, int rows) {
    a.ntiles = (a.ny + rows - 1) / rows;
    if (a.yhi > a.ylo) {
        a.tile0 = a.ylo / rows;
        a.nband = (a.yhi + rows - 1) / rows - a.tile0;
    } else {
        a.tile0 = 0;
        a.nband = 0;
    }
    return (unsigned)(a.nband * (a.R + 1) + (a.ntiles - a.nband));
}

template <typename T, int N>
struct DancerRowLaunch {
    template <int XB>
    static void launch_rows(const DancerRowArgs<T>& a, dim3 grid, dim3 block, size_t smem, hipStream_t s) {
        if constexpr (XB > 0 && N < 128) {
            launch_rows<0>(a, grid, block, smem, s);
        } else if constexpr (std::is_same<T, float>::value) {
            allow_lds((const void*)k_dancer_rows_reg4<N, XB>, smem);
            hipLaunchKernelGGL((k_dancer_rows_reg4<N, XB>), grid, block, smem, s, a);
        } else {
            allow_lds((const void*)k_dancer_rows_reg<T, N, XB>, smem);
            hipLaunchKernelGGL((k_dancer_rows_reg<T, N, XB>), grid, block, smem, s, a);
        }
    }
    static void run(DancerRowArgs<T> a, int fb, int rows_req, hipStream_t s) {
        if constexpr (N >= 16) {
            if (use_reg_fft()) {
                using C = RegE<T, N>;
                int rows = rows_req > 0 ? rows_req : std::max(1, 128 / C::TT);
                rows = std::max(1, std::min(rows, 256 / C::TT));
                const size_t smem = sizeof(cplx<typename Storage<T>::type>) * (size_t)C::LD * (2 * rows);
                const unsigned nblk = dancer_tiles(a, rows);
                const dim3 grid(nblk, fb), block(2 * rows * C::TT);
                const int xb = (a.keep_tab && (a.flags & DF_FWD)) ? a.keep_band : 0;
                if (xb == 5) {
                    launch_rows<5>(a, grid, block, smem, s);
                } else if (xb == 3) {
                    launch_rows<3>(a, grid, block, smem, s);
                } else {
                    launch_rows<0>(a, grid, block, smem, s);
                }
                BF_CHECK_LAUNCH();
                return;
            }
        }
        const int TPL = LineCfg<N>::TPL;
        const int n = N >= 16 ? N : a.nx;
        int rows = rows_req > 0 ? rows_req : std::max(1, 128 / TPL);
        rows = std::max(1, std::min(rows, 512 / TPL));
        const size_t smem = sizeof(cplx<typename Compute<T>::type>) * (size_t)n * (2 * rows) * 2;
        const unsigned nblk = dancer_tiles(a, rows);
        allow_lds((const void*)k_dancer_rows<T, N>, smem);
        hipLaunchKernelGGL((k_dancer_rows<T, N>), dim3(nblk, fb), dim3(2 * rows * TPL), smem, s, a);
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
    BF_TYPES(T);
    bf_ctx* ctx;
    const Grid* g;
    int kind;
    double R2;
    const cplx<C>* twx;
    const cplx<C>* twy;
    const zd* surfaces;  // device, may be null
    DevBuf ktab, rtab, kdesc, rdesc, scratch;
    int ntab = 0;              // k-space tables built
    std::vector<int> ktab_of;  // per region entry: k-space table index
    std::vector<int> rtab_of;  // per region entry: rtab index or -1
    int col_tx = 0;            // autotuned column tile (0: default)
    // Quarter k-space tables (see k_build_ktab): set for complex64 fields on
    // power-of-two grids; only the register propagation pass reads them.
    bool quarter = false;
    // aperture row band [blo, bhi) of propagate() (bhi <= blo: every row)
    int blo = 0, bhi = 0;

    PropEngine(bf_ctx* c, const Grid* gr, int k, double diskR, const zd* surf)
        : ctx(c), g(gr), kind(k), R2(diskR * diskR), surfaces(surf) {
        twx = twiddle_table<C>(ctx, g->nx);
        twy = twiddle_table<C>(ctx, g->ny);
    }

    // Build the k-space and real-space tables for a list of regions.  With
    // dedup, regions whose k-space tables would be bit-identical (same k0^2,
    // dz and layout, e.g. every disk of a booster) share one table; ktab_of
    // maps region entries to tables either way.
    void build_tables(const std::vector<RegionParams>& regs, bool dedup = false) {
        const size_t K = g->K();
        const int nreg = (int)regs.size();
        std::vector<KTabDesc> kd;
        std::vector<RTabDesc> rd;
        rtab_of.assign(nreg, -1);
        ktab_of.assign(nreg, -1);
        std::map<std::array<uint64_t, 4>, int> seen;
        for (int i = 0; i < nreg; i++) {
            const RegionParams& p = regs[i];
            zd eps = p.eps;
            if (kind == BF_PROP_PROPAGATOR && eps.im == 0.0) eps.im = -0.0;  // src/boost_fractor.jl:140
            const zd k0 = k0_of(eps, p.lambda);
            KTabDesc d;
            d.k0sq = cmul(k0, k0);
            d.dz = p.dz;
            d.shifted = (kind == BF_PROP_MOMENTUM) ? 1 : 0;
            d.pad = 0;
            std::array<uint64_t, 4> key;
            std::memcpy(&key[0], &d.k0sq.re, 8);
            std::memcpy(&key[1], &d.k0sq.im, 8);
            std::memcpy(&key[2], &d.dz, 8);
            key[3] = (uint64_t)d.shifted;
            auto it = dedup ? seen.find(key) : seen.end();
            if (it != seen.end()) {
                ktab_of[i] = it->second;
            } else {
                ktab_of[i] = (int)kd.size();
                if (dedup) seen.emplace(key, ktab_of[i]);
                kd.push_back(d);
            }
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
        ntab = (int)kd.size();
        if (kind == BF_PROP_1D) return;
        const double scale = 1.0 / (double)((int64_t)g->nx * g->ny);  // ifft normalisation folded in
        const size_t KT = quarter ? (size_t)(g->nx / 2) * (g->ny / 2) : K;
        if (ktab.bytes < sizeof(cplx<C>) * KT * ntab) ktab.alloc(sizeof(cplx<C>) * KT * ntab);
        if (kdesc.bytes < sizeof(KTabDesc) * ntab) kdesc.alloc(sizeof(KTabDesc) * ntab);
        upload(kdesc.p, kd.data(), sizeof(KTabDesc) * ntab, ctx->stream);
        hipLaunchKernelGGL(k_build_ktab<C>, dim3(grid1d(KT, 256, 1024), ntab), dim3(256), 0, ctx->stream,
                           ktab.as<cplx<C>>(), kdesc.as<KTabDesc>(), g->nx, g->ny, g->kX.as<double>(),
                           g->kY.as<double>(), scale, quarter ? 1 : 0);
        BF_CHECK_LAUNCH();
        if (!rd.empty()) {
            if (rtab.bytes < sizeof(cplx<C>) * K * rd.size()) rtab.alloc(sizeof(cplx<C>) * K * rd.size());
            if (rdesc.bytes < sizeof(RTabDesc) * rd.size()) rdesc.alloc(sizeof(RTabDesc) * rd.size());
            upload(rdesc.p, rd.data(), sizeof(RTabDesc) * rd.size(), ctx->stream);
            hipLaunchKernelGGL(k_build_rtab<C>, dim3(grid1d(K, 256, 1024), (unsigned)rd.size()), dim3(256), 0,
                               ctx->stream, rtab.as<cplx<C>>(), rdesc.as<RTabDesc>(), g->nx, g->ny, g->X.as<double>(),
                               g->Y.as<double>(), surfaces);
            BF_CHECK_LAUNCH();
        }
    }

    bool has_rtab() const { return rtab.p != nullptr && !rtab_of.empty() &&
                                   std::any_of(rtab_of.begin(), rtab_of.end(), [](int v) { return v >= 0; }); }

    // ylo/yhi: only rows in [ylo, yhi) when yhi > ylo
    void rows(cplx<S>* base, const FieldDesc* fd, int nf, uint32_t flags, const int* active, int ylo = 0,
              int yhi = 0) {
        RowArgs<T> a;
        a.base = base; a.fd = fd; a.nf = nf; a.nx = g->nx; a.ny = g->ny; a.tw = twx; a.flags = flags;
        a.rtab = rtab.as<cplx<C>>(); a.X = g->X.as<double>(); a.Y = g->Y.as<double>(); a.R2 = R2; a.active = active;
        a.ylo = ylo; a.yhi = yhi;
        dispatch_n<RowLaunch, T>(g->nx, a, ctx->stream);
    }
    // ylo/yhi: row band (yhi < 0: all rows); all_whole: store every row of every field
    void cols(const cplx<S>* src, cplx<S>* dst, const FieldDesc* fd, int nf, uint32_t flags, const int* active,
              int ylo = 0, int yhi = -1, bool all_whole = false) {
        ColArgs<T> a;
        a.src = src; a.dst = dst; a.fd = fd; a.nf = nf; a.nx = g->nx; a.ny = g->ny; a.tw = twy; a.flags = flags;
        a.ktab = ktab.as<cplx<C>>(); a.tx = 0; a.active = active;
        a.kq = quarter ? g->nx / 2 : 0;
        a.all_whole = all_whole ? 1 : 0;
        a.ylo = ylo; a.yhi = yhi < 0 ? g->ny : yhi;
        dispatch_n<ColLaunch, T>(g->ny, a, col_tx, ctx->stream);
    }

    // Full propagation of nf fields (field j uses tables fd[j].ktab / fd[j].rtab).
    // `scr` must hold as many elements as `base` spans (momentum-space kind only).
    // the masked forward half of a propagation (propagator and propagatorNoTilts):
    // row FFT of the masked fields and the column pass; the inverse row pass follows.
    // band_out: the column pass stores the band rows only (the caller reads no other)
    void propagate_head(cplx<S>* base, const FieldDesc* fd, int nf, const int* active, bool band_out = false) {
        if (bhi > blo) {
            // the mask zeroes every row outside the aperture band: the masked row
            // FFT skips them and the column pass reads only the band, storing
            // every row of the propagated field (or the band only)
            rows(base, fd, nf, RF_MASK | RF_FWD, active, blo, bhi);
            cols(base, base, fd, nf, CF_FWD | CF_KMID | CF_INV, active, blo, bhi, !band_out);
        } else {
            rows(base, fd, nf, RF_MASK | RF_FWD, active);
            cols(base, base, fd, nf, CF_FWD | CF_KMID | CF_INV, active);
        }
    }

    void propagate(cplx<S>* base, const FieldDesc* fd, int nf, const zd* ph1d, cplx<S>* scr, const int* active) {
        const size_t K = g->K();
        switch (kind) {
            case BF_PROP_PROPAGATOR:
            case BF_PROP_NOTILTS:
                propagate_head(base, fd, nf, active);
                rows(base, fd, nf, RF_INV | (kind == BF_PROP_PROPAGATOR ? RF_RTAB : 0u), active);
                break;
            case BF_PROP_MOMENTUM:
                cols(base, scr, fd, nf, CF_INSHIFT | CF_KLOAD | CF_INV, active);
                rows(scr, fd, nf, RF_INV | RF_MASK | RF_RTAB | RF_FWD, active);
                cols(scr, base, fd, nf, CF_FWD | CF_OUTSHIFT, active);
                break;
            default:
                hipLaunchKernelGGL(k_scale_fields<T>, dim3(grid1d(K, 256, 256), nf), dim3(256), 0, ctx->stream, (S*)base,
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
                     double* bfield, double velocity_x, double f, double* out) {
    hipStream_t s = ctx->stream;
    Grid g;
    g.upload(coords, s);
    const size_t K = g.K();
    std::vector<zd> ax = dance_intro_amplitudes(nd, reinterpret_cast<const zd*>(eps), r1);
    DevBuf dax(sizeof(zd) * ax.size()), dout(sizeof(zd) * K * 2 * nd), db;
    upload(dax.p, ax.data(), sizeof(zd) * ax.size(), s);
    if (bfield) {
        const size_t nb = K * (size_t)(nd + 1);
        db.alloc(sizeof(zd) * nb);
        upload(db.p, bfield, sizeof(zd) * nb, s);
        if (velocity_x != 0.0) {
            const double ma = 2.0 * M_PI / (299792458.0 / f);
            zd a = cmul(cmk<double>(0.0, -1.0), cmk<double>(ma, 0.0));
            a = cmk<double>(a.re * -velocity_x, a.im * -velocity_x);
            hipLaunchKernelGGL(k_bfield_velocity, dim3(grid1d(nb, 256, 2048)), dim3(256), 0, s, db.as<zd>(), nb, K,
                               g.nx, g.X.as<double>(), a);
            BF_CHECK_LAUNCH();
            download(bfield, db.p, sizeof(zd) * nb, s);  // the reference multiplies the caller's array in place
        }
    }
    hipLaunchKernelGGL(k_dance_intro<double>, dim3(grid1d(K, 256, 512), 2 * nd, 1), dim3(256), 0, s,
                       dout.as<zd>(), nd, K, g.nx, dax.as<zd>(), g.X.as<double>(), g.Y.as<double>(), diskR * diskR,
                       (size_t)0, bfield ? db.as<zd>() : nullptr);
    BF_CHECK_LAUNCH();
    download(out, dout.p, sizeof(zd) * K * 2 * nd, s);
    BF_HIP(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------------------
// dancer plan
// ---------------------------------------------------------------------------
// Quarter k-space tables for complex64 fields: the phase of a mirrored bin comes
// from its mirror's k, equal up to the rounding of the k range (~1e-15 rad per
// propagation), far inside what complex64 storage rounds away; the FP64 path
// keeps full tables.  Needs the register propagation pass (square power-of-two
// grid) and a mirror-symmetric k range.  Option qtab=0 switches it off.
template <typename T>
static bool quarter_tables_ok(const bf_coords* co, int kind) {
    const int n = co->nx;
    auto mirrored = [](const double* k, int len) {
        double mx = 0.0;
        for (int j = 0; j < len; j++) mx = std::max(mx, std::fabs(k[j]));
        for (int j = 0; j < len; j++)
            if (!(std::fabs(k[j] + k[len - 1 - j]) <= 1e-12 * mx)) return false;
        return true;
    };
    return !std::is_same<T, double>::value && opts().qtab && use_reg_fft() &&
           (kind == BF_PROP_PROPAGATOR || kind == BF_PROP_NOTILTS) && co->nx == co->ny && n >= 16 && n <= 1024 &&
           (n & (n - 1)) == 0 && mirrored(co->kX, co->nx) && mirrored(co->kY, co->ny);
}

// Rows y with Y[y]^2 + min_x X[x]^2 < diskR^2 (a relative margin of 1e-9 keeps
// every row the device mask could keep): [lo, hi), or [0, 0) when none.
static void aperture_band(const bf_coords* co, double diskR, int& lo, int& hi) {
    double minx2 = INFINITY;
    for (int x = 0; x < co->nx; x++) minx2 = std::min(minx2, co->X[x] * co->X[x]);
    const double lim = diskR * diskR * (1.0 + 1e-9);
    lo = co->ny;
    hi = 0;
    for (int y = 0; y < co->ny; y++)
        if (co->Y[y] * co->Y[y] + minx2 < lim) {
            lo = std::min(lo, y);
            hi = y + 1;
        }
    if (lo >= hi) lo = hi = 0;
}

template <typename T>
struct DancerPlan : DancerPlanBase {
    BF_TYPES(T);
    // The fused bounce loop of one frequency batch (2 (nmax+1) launches) captured
    // once and replayed by later runs.  Every buffer its kernels read belongs to the
    // plan and keeps its address, except the engine's phase/tilt tables, which grow
    // by reallocation: a graph is recaptured when their addresses change.
    struct BatchGraph {
        hipGraphExec_t exec = nullptr;
        const void* ktab = nullptr;
        const void* rtab = nullptr;
    };
    std::vector<BatchGraph> graphs;
    Grid g;
    int kind, R, nmax, nfreq, fb, K_;
    double amin, diskR;
    std::vector<double> freqs;
    std::vector<zd> eps;  // trimmed: regions 1..nreg-1 of the setup
    std::vector<double> dist, tx, ty;
    bool has_surf = false;
    DevBuf surf, coef, fields, scratch, Eout, hist, power, active, amax, count, fdesc, fdesc_col, ridx, ax, emit,
        reflect, ph1d, keep_tab;
    int keep_band = 0;  // band class of keep_tab (see prune_class)
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
        for (auto& bg : graphs)
            if (bg.exec) (void)hipGraphExecDestroy(bg.exec);
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
        // Row band of the aperture: rows y with Y[y]^2 + min_x X[x]^2 < diskR^2
        // (a relative margin of 1e-9 keeps every row the device mask could keep).
        // The mask zeroes every other row before each propagation, so the fused
        // bounce skips them (except for the exit field's Eout).  The amin
        // criterion reads every row, so it keeps the full range.
        band_lo = 0;
        band_hi = g.ny;
        if (opt.band && amin == 0.0 && (kind == BF_PROP_PROPAGATOR || kind == BF_PROP_NOTILTS))
            aperture_band(co, diskR, band_lo, band_hi);
        const bool qtab = quarter_tables_ok<T>(co, kind);
        // Frequency batch: what one bounce of a batch touches (the band rows of every
        // field, plus one k-space table per distinct (eps, dz)) is kept near
        // 192 MB, the size of the Infinity Cache; the frequencies are then spread
        // evenly over the batches.  Measured on the 20-disk booster at 256^2:
        // 5 frequencies (complex64 fields) and 4 (complex128) are the fastest.
        const size_t per = sizeof(cplx<S>) * K * 2 * R;
        if (opt.freq_batch > 0) {
            fb = opt.freq_batch;
        } else {
            std::vector<std::array<uint64_t, 3>> keys((size_t)R);  // bit patterns: NaN-safe ordering
            for (int i = 0; i < R; i++) {
                std::memcpy(&keys[i][0], &eps[i].re, 8);
                std::memcpy(&keys[i][1], &eps[i].im, 8);
                std::memcpy(&keys[i][2], &dist[i], 8);
            }
            std::sort(keys.begin(), keys.end());
            const size_t ntab = (size_t)(std::unique(keys.begin(), keys.end()) - keys.begin());
            const size_t touched = sizeof(cplx<S>) * (size_t)(band_hi - band_lo) * g.nx * 2 * R +
                                   ntab * (qtab ? K / 4 : K) * sizeof(cplx<C>);
            const size_t budget = (size_t)192 << 20;
            fb = (int)std::max<size_t>(1, budget / std::max<size_t>(touched, 1));
        }
        fb = std::max(1, std::min(fb, std::min(nfreq, 64)));
        if (opt.freq_batch <= 0) {
            const int nb = (nfreq + fb - 1) / fb;
            fb = (nfreq + nb - 1) / nb;
        }
        freq_batch = fb;
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
        fdesc_col.alloc(2 * sizeof(FieldDesc) * fd.size());
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
        pe->quarter = qtab;
        // aperture bits of the register row pass (thread t of a line holds x = t + (nx/16) m, m < 16)
        if (g.nx >= 16 && g.nx % 16 == 0) {
            const int TT = g.nx / 16;
            const double R2 = diskR * diskR;
            std::vector<uint32_t> kb((size_t)g.ny * TT);
            for (int y = 0; y < g.ny; y++)
                for (int t = 0; t < TT; t++) {
                    uint32_t bits = 0;
                    for (int m = 0; m < 16; m++) bits |= (in_aperture(co->X[t + TT * m], co->Y[y], R2) ? 1u : 0u) << m;
                    kb[(size_t)y * TT + t] = bits;
                }
            uint32_t any = 0;
            for (uint32_t w : kb) any |= w;
            keep_band = any ? prune_class(__builtin_ctz(any), 31 - __builtin_clz(any), 16) : 0;
            keep_tab.alloc(sizeof(uint32_t) * kb.size());
            upload(keep_tab.p, kb.data(), sizeof(uint32_t) * kb.size(), s);
        }
        BF_HIP(hipStreamSynchronize(s));
    }

    // The fused bounces of one frequency batch: column pass, row pass (boundary
    // update, Eout, next aperture) and the amin test, nmax + 1 times.  Only kernel
    // launches (no allocation, copy or synchronisation), so a stream can capture it.
    void bounce_loop(int f0, int nb, zd* eo_target) {
        hipStream_t s = ctx->stream;
        const size_t K = g.K();
        int rm = 0;
        for (int n = 0; n <= nmax; n++) {
            tmark(n, 1, 0);
            // every slot of the table-ordered list (inactive slots return at once)
            const int ncol = (int)fd_host.size();
            pe->cols(fields.as<cplx<S>>(), fields.as<cplx<S>>(), fdesc_col.as<FieldDesc>() + (size_t)rm * ncol,
                     ncol, CF_FWD | CF_KMID | CF_INV, active.as<int>(), band_lo, band_hi);
            tmark(n, 1, 1);
            DancerRowArgs<T> a;
            a.fields = fields.as<cplx<S>>();
            a.R = R; a.nx = g.nx; a.ny = g.ny; a.rm = rm; a.tw = pe->twx;
            a.rtab = pe->has_rtab() ? pe->rtab.template as<cplx<C>>() : nullptr;
            a.rtab_idx = ridx.as<int>();
            a.coef = coef.as<zd>();
            a.X = g.X.as<double>(); a.Y = g.Y.as<double>(); a.R2 = diskR * diskR;
            a.keep_tab = keep_tab.as<uint32_t>();
            a.keep_band = keep_band;
            a.flags = (n < nmax ? DF_FWD : 0u) | (amin != 0.0 ? DF_AMAX : 0u) |
                      (nslices > 1 ? DF_EOUT_ASSIGN : 0u);
            a.Eout = eo_target;
            a.eout_fstride = K * nslices;
            a.slice = nslices > 1 ? n : 0;
            a.amax = amax.as<double>();
            a.active = active.as<int>();
            a.count = count.as<int>() + f0;
            a.ylo = band_lo;
            a.yhi = band_hi;
            tmark(n, 0, 0);
            dispatch_n<DancerRowLaunch, T>(g.nx, a, nb, opt.rows, s);
            tmark(n, 0, 1);
            if (amin != 0.0) {
                launch_amin_check(s, active.as<int>(), amax.as<double>(), amin, nb);
                BF_CHECK_LAUNCH();
            }
            rm = 1 - rm;
        }
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
            pe->build_tables(regs, true);
            std::vector<int> ri((size_t)fb * R);
            for (size_t q = 0; q < ri.size(); q++) ri[q] = pe->rtab_of.empty() ? -1 : pe->rtab_of[q];
            upload(ridx.p, ri.data(), sizeof(int) * ri.size(), s);
            // field (f, region, slot) -> shared k-space table; the unfused propagators
            // also read the tilt/surface table index from the descriptors
            for (int f = 0; f < fb; f++)
                for (int q = 0; q < 2 * R; q++) {
                    FieldDesc& d = fd_host[(size_t)f * 2 * R + q];
                    d.ktab = pe->ktab_of[(size_t)f * R + q / 2];
                    d.rtab = fused ? -1 : ri[(size_t)f * R + q / 2];
                }
            upload(fdesc.p, fd_host.data(), sizeof(FieldDesc) * fd_host.size(), s);
            if (fused) {
                // Column-pass order: fields grouped by k-space table, so the blocks that
                // read one table run back to back and find it in L2 (every disk of a
                // booster shares a table).  Two copies: the exit field (region R-1,
                // right-moving slot rm), needed whole for Eout, is flagged per slot.
                const size_t nfd = fd_host.size();
                std::vector<size_t> ord(nfd);
                for (size_t q = 0; q < nfd; q++) ord[q] = q;
                std::stable_sort(ord.begin(), ord.end(),
                                 [&](size_t x, size_t y) { return fd_host[x].ktab < fd_host[y].ktab; });
                std::vector<FieldDesc> col(2 * nfd);
                for (int r = 0; r < 2; r++)
                    for (size_t q = 0; q < nfd; q++) {
                        FieldDesc d = fd_host[ord[q]];
                        d.pad = ((int)(ord[q] % (2 * R)) == 2 * (R - 1) + r) ? FD_WHOLE : 0;
                        col[(size_t)r * nfd + q] = d;
                    }
                upload(fdesc_col.p, col.data(), sizeof(FieldDesc) * col.size(), s);
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
                                   fields.as<cplx<S>>(), emit.as<zd>(), fstride, fstride);
            } else if (opt.reflect) {
                BF_HIP(hipMemsetAsync(fields.p, 0, sizeof(cplx<S>) * fstride * nb, s));
                // fields[R-1][leftmoving = slot 1] = reflect
                hipLaunchKernelGGL(k_broadcast<T>, dim3(grid1d(K, 256, 1024), nb), dim3(256), 0, s,
                                   fields.as<cplx<S>>() + ((size_t)(R - 1) * 2 + 1) * K, reflect.as<zd>(), K, fstride);
            } else {
                hipLaunchKernelGGL(k_dance_intro<T>, dim3(grid1d(K, 256, 512), 2 * R, nb), dim3(256), 0, s,
                                   fields.as<cplx<S>>(), R, K, g.nx, ax.as<zd>(), g.X.as<double>(), g.Y.as<double>(),
                                   diskR * diskR, fstride, (const zd*)nullptr);
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
                pe->rows(fields.as<cplx<S>>(), fd, nf, RF_MASK | RF_FWD, active.as<int>());
                if (opt.graph && !timing && nmax >= 0) {
                    const size_t bi = (size_t)(f0 / fb);
                    if (graphs.size() <= bi) graphs.resize(bi + 1);
                    BatchGraph& bg = graphs[bi];
                    if (bg.exec && (bg.ktab != pe->ktab.p || bg.rtab != pe->rtab.p)) {
                        BF_HIP(hipGraphExecDestroy(bg.exec));
                        bg.exec = nullptr;
                    }
                    if (!bg.exec) {
                        BF_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                        hipGraph_t gr = nullptr;
                        try {
                            bounce_loop(f0, nb, eo_target);
                        } catch (...) {
                            (void)hipStreamEndCapture(s, &gr);
                            if (gr) (void)hipGraphDestroy(gr);
                            throw;
                        }
                        BF_HIP(hipStreamEndCapture(s, &gr));
                        const hipError_t e = hipGraphInstantiate(&bg.exec, gr, nullptr, nullptr, 0);
                        (void)hipGraphDestroy(gr);
                        BF_HIP(e);
                        bg.ktab = pe->ktab.p;
                        bg.rtab = pe->rtab.p;
                    }
                    BF_HIP(hipGraphLaunch(bg.exec, s));
                } else {
                    bounce_loop(f0, nb, eo_target);
                    tcollect(nmax + 1);
                }
            } else {
                for (int n = 0; n <= nmax; n++) {
                    pe->propagate(fields.as<cplx<S>>(), fd, nf, ph1d.as<zd>(), scratch.as<cplx<S>>(),
                                  active.as<int>());
                    hipLaunchKernelGGL(k_dancer_update<T>, dim3(grid1d(K, 256, 256), R + 1, nb), dim3(256), 0, s,
                                       fields.as<cplx<S>>(), R, K, rm, coef.as<zd>(), eo_target, K * nslices,
                                       nslices > 1 ? n : 0, nslices > 1 ? 1 : 0,
                                       amin != 0.0 ? amax.as<double>() : nullptr, active.as<int>(),
                                       count.as<int>() + f0);
                    BF_CHECK_LAUNCH();
                    if (amin != 0.0) {
                        launch_amin_check(s, active.as<int>(), amax.as<double>(), amin, nb);
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

static DancerPlanBase* make_dancer_plan_fixed(bf_ctx* ctx, const bf_setup* st, const bf_coords* co, double amin,
                                              int nmax, const double* fr, int nfr, int kind, double diskR,
                                              int precision, const DancerOptions& o) {
    if (precision == BF_PREC_MIXED) {
        auto* p = new DancerPlan<float>();
        try { p->init(ctx, st, co, amin, nmax, fr, nfr, kind, diskR, o); } catch (...) { delete p; throw; }
        return p;
    }
    if (precision == BF_PREC_FP32) {
        auto* p = new DancerPlan<f32pure>();
        try { p->init(ctx, st, co, amin, nmax, fr, nfr, kind, diskR, o); } catch (...) { delete p; throw; }
        return p;
    }
    auto* p = new DancerPlan<double>();
    try { p->init(ctx, st, co, amin, nmax, fr, nfr, kind, diskR, o); } catch (...) { delete p; throw; }
    return p;
}

// Autotuner of the dancer's launch shapes (a measured plan, where the reference's
// FFTW.fft! calls, src/boost_fractor.jl:163-164, plan by estimate): short runs of the plan's own inputs
// (nmax cut to TUNE_NMAX, the frequencies cut to a few batches) time candidate
// frequency batches around the cache-size heuristic, then column-pass tiles,
// then row-pass rows per block (coordinate descent), per shape key.
static std::string dancer_tune_key(const bf_coords* co, const bf_setup* st, int precision, int kind, double amin,
                                   int nfr, const DancerOptions& o) {
    return "dancer p" + std::to_string(precision) + " " + std::to_string(co->nx) + "x" + std::to_string(co->ny) +
           " reg" + std::to_string(st->nreg) + " k" + std::to_string(kind) + " a" + std::to_string(amin != 0.0) +
           " f" + std::to_string(std::min(nfr, 64)) + " s" + std::to_string(o.returnsum) +
           std::to_string(o.immediatesum) + " e" + std::to_string(o.emit != nullptr) +
           std::to_string(o.reflect != nullptr);
}

static void tune_dancer(bf_ctx* ctx, const bf_setup* st, const bf_coords* co, double amin, int nmax,
                        const double* fr, int nfr, int kind, double diskR, int precision, const DancerOptions& o,
                        const std::string& key) {
    constexpr int TUNE_NMAX = 20;
    const auto t_start = std::chrono::steady_clock::now();
    hipStream_t s = ctx->stream;
    DancerOptions base = o;
    base.graph = 0;
    base.returnsum = 1;
    base.immediatesum = 1;
    const int nm = std::min(nmax, TUNE_NMAX);
    // the heuristic batch, from a plan that picks it itself
    int fb0;
    {
        std::unique_ptr<DancerPlanBase> p(make_dancer_plan_fixed(ctx, st, co, amin, nm, fr, nfr, kind, diskR,
                                                                 precision, base));
        fb0 = p->freq_batch;
    }
    const int fbmax = std::min(nfr, 64);
    const int nprobe = std::min(nfr, 2 * std::min(fbmax, 2 * fb0));  // every candidate runs the same points
    auto time_of = [&](const DancerOptions& t) {
        std::unique_ptr<DancerPlanBase> p(make_dancer_plan_fixed(ctx, st, co, amin, nm, fr, nprobe, kind, diskR,
                                                                 precision, t));
        p->run();  // first run: lazy tables
        BF_HIP(hipStreamSynchronize(s));
        const auto t0 = std::chrono::steady_clock::now();
        p->run();
        BF_HIP(hipStreamSynchronize(s));
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    };
    DancerOptions best = base;
    best.freq_batch = fb0;
    double tbest = time_of(best);
    for (int fb : {std::max(1, fb0 * 2 / 3), fb0 * 3 / 2, fb0 * 2}) {
        if (fb < 1 || fb > fbmax || fb == best.freq_batch) continue;
        DancerOptions t = best;
        t.freq_batch = fb;
        const double tt = time_of(t);
        if (tt < tbest) { tbest = tt; best = t; }
    }
    // column tile (columns per column-pass block; the kernel's default is 128 bytes of a line) and rows
    // per row-pass block (default 128 threads): halved and doubled, in the kernels' valid ranges
    const int esize = precision == BF_PREC_FP64 ? 16 : 8;
    const int tx0 = 128 / esize, TT = co->nx >= 16 ? co->nx / 16 : 1;
    for (int tx : {tx0 / 2, tx0 * 2}) {
        if (tx < 1 || tx > 512 / std::max(1, co->ny / 16)) continue;
        DancerOptions t = best;
        t.col_tx = tx;
        const double tt = time_of(t);
        if (tt < tbest) { tbest = tt; best = t; }
    }
    const int rows0 = std::max(1, 128 / TT);
    for (int r : {rows0 / 2, rows0 * 2}) {
        if (r < 1 || r > 256 / TT) continue;
        DancerOptions t = best;
        t.rows = r;
        const double tt = time_of(t);
        if (tt < tbest) { tbest = tt; best = t; }
    }
    ctx->tuned[key] = {best.freq_batch, best.col_tx, best.rows};
    ctx->tune_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    save_tuned(ctx, key);
}

DancerPlanBase* make_dancer_plan(bf_ctx* ctx, const bf_setup* st, const bf_coords* co, double amin, int nmax,
                                 const double* fr, int nfr, int kind, double diskR, int precision,
                                 const DancerOptions& o) {
    if (opts().autotune && o.freq_batch == 0 && o.col_tx == 0 && o.rows == 0 && nmax >= 1) {
        const std::string key = dancer_tune_key(co, st, precision, kind, amin, nfr, o);
        if (!ctx->tuned.count(key)) tune_dancer(ctx, st, co, amin, nmax, fr, nfr, kind, diskR, precision, o, key);
        const std::vector<int64_t>& v = ctx->tuned[key];
        if (v.size() == 3) {
            DancerOptions t = o;
            const int nb = (int)((nfr + v[0] - 1) / v[0]);  // the tuned batch, spread evenly over the batches
            t.freq_batch = (nfr + nb - 1) / nb;
            t.col_tx = (int)v[1];
            t.rows = (int)v[2];
            return make_dancer_plan_fixed(ctx, st, co, amin, nmax, fr, nfr, kind, diskR, precision, t);
        }
    }
    return make_dancer_plan_fixed(ctx, st, co, amin, nmax, fr, nfr, kind, diskR, precision, o);
}

// ---------------------------------------------------------------------------
// cheerleader (src/boost_fractor_recurprop.jl:226-363), Gauss-Seidel sweep,
// batched over frequencies.  fields[f][Nd][K].
// ---------------------------------------------------------------------------
// F[i+1] += t*F[i]; F[i] *= r      (prologue and last region, :285-286, :343-344)
template <typename T>
__global__ void k_ch_push(cplx<typename Storage<T>::type>* F, int Nd, size_t K, int i, zd t_, zd r_,
                          const int* active) {
    BF_TYPES(T);
    const int f = blockIdx.y;
    if (active && !active[f]) return;
    const cplx<C> t = cmk<C>((C)t_.re, (C)t_.im), r = cmk<C>((C)r_.re, (C)r_.im);
    cplx<S>* Fi = F + ((size_t)f * Nd + i) * K;
    cplx<S>* Fn = Fi + K;
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < K; p += (size_t)gridDim.x * blockDim.x) {
        const cplx<C> v = ccast<C>(Fi[p]);
        Fn[p] = ccast<S>(cadd(ccast<C>(Fn[p]), cmul(t, v)));
        Fi[p] = ccast<S>(cmul(v, r));
    }
}
// F[0] += tl*F[1]; F[1] *= rl    (first region, :318-320)
template <typename T>
__global__ void k_ch_first(cplx<typename Storage<T>::type>* F, int Nd, size_t K, zd tl_, zd rl_, const int* active) {
    BF_TYPES(T);
    const int f = blockIdx.y;
    if (active && !active[f]) return;
    const cplx<C> tl = cmk<C>((C)tl_.re, (C)tl_.im), rl = cmk<C>((C)rl_.re, (C)rl_.im);
    cplx<S>* F0 = F + (size_t)f * Nd * K;
    cplx<S>* F1 = F0 + K;
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < K; p += (size_t)gridDim.x * blockDim.x) {
        const cplx<C> v = ccast<C>(F1[p]);
        F0[p] = ccast<S>(cadd(ccast<C>(F0[p]), cmul(tl, v)));
        F1[p] = ccast<S>(cmul(v, rl));
    }
}
// pair update at boundary i (:330-334)
template <typename T>
__global__ void k_ch_pair(cplx<typename Storage<T>::type>* F, int Nd, size_t K, int i, zd rl_, zd tr_, zd rr_, zd tl_,
                          const int* active) {
    BF_TYPES(T);
    const int f = blockIdx.y;
    if (active && !active[f]) return;
    const cplx<C> rl = cmk<C>((C)rl_.re, (C)rl_.im), tr = cmk<C>((C)tr_.re, (C)tr_.im);
    const cplx<C> rr = cmk<C>((C)rr_.re, (C)rr_.im), tl = cmk<C>((C)tl_.re, (C)tl_.im);
    cplx<S>* Fi = F + ((size_t)f * Nd + i) * K;
    cplx<S>* Fn = Fi + K;
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < K; p += (size_t)gridDim.x * blockDim.x) {
        const cplx<C> a = ccast<C>(Fi[p]), c = ccast<C>(Fn[p]);
        Fn[p] = ccast<S>(cadd(cmul(c, rl), cmul(tr, a)));
        Fi[p] = ccast<S>(cadd(cmul(a, rr), cmul(tl, c)));
    }
}
// The pair step of the cheerleader in one pass over rows: the propagation's
// last row pass (inverse FFT along x, tilt/surface table) for the region pair
// (i, i+1) and then the boundary update of k_ch_pair on the same row.  Row y is
// two lines of a block (field i, field i+1) that exchange values through LDS;
// both values are rounded to the storage precision before the update, as the
// separate passes round them in HBM.  Rows [ylo, yhi) only: the whole grid, or
// the aperture band when nothing reads the pair's other rows before the mask
// of their next propagation zeroes them.  With xout the kernel also starts
// that next propagation: the updated rows are masked by the aperture and
// transformed along x (the masked row FFT of k_rows_reg) before they are stored.
template <typename T>
struct ChPairArgs {
    cplx<typename Storage<T>::type>* F;  // [fb][Nd][K]
    // entry e (one frequency of one region pair (i, i+1)): fd[2e] field i,
    // fd[2e+1] field i+1; i = fd[2e].pad >> FD_PAIR_SHIFT
    const FieldDesc* fd;
    int ny, rows;
    const cplx<typename Compute<T>::type>* tw;
    const cplx<typename Compute<T>::type>* rtab;  // null: no tilt/surface tables
    const zd* coef;  // [i][4]: rl, tr, rr, tl of boundary i
    const int* active;
    int ylo, yhi;  // rows processed
    int xout;      // store the masked row transform of the updated fields
    const double* X;
    const double* Y;
    double R2;  // aperture radius squared
};

template <typename T, int N>
__global__ void __launch_bounds__(512) k_ch_pair_rows(ChPairArgs<T> a) {
    BF_TYPES(T);
    extern __shared__ unsigned char smem_raw[];
    constexpr int E = RegE<T, N>::E, TT = RegE<T, N>::TT, LD = RegE<T, N>::LD;
    cplx<S>* smem = reinterpret_cast<cplx<S>*>(smem_raw);
    const int e = blockIdx.y;
    const FieldDesc d0 = a.fd[2 * e];
    if (a.active && !a.active[d0.group]) return;  // uniform over the block
    const int li = threadIdx.x / TT, t = threadIdx.x % TT;
    const int side = li & 1;  // 0: field i, 1: field i+1
    const int y = a.ylo + blockIdx.x * a.rows + (li >> 1);
    const bool ok = y < a.yhi;
    const int yc = ok ? y : a.yhi - 1;
    const size_t K = (size_t)N * a.ny;
    cplx<S>* s = smem + (size_t)li * LD;
    const cplx<S>* sp = smem + (size_t)(li ^ 1) * LD;  // partner line, same y
    const FieldDesc d = a.fd[2 * e + side];
    cplx<S>* line = a.F + d.off + (size_t)yc * N + t;
    cplx<C> v[E];
    {
        cplx<S> raw[E];
#pragma unroll
        for (int m = 0; m < E; m++) raw[m] = line[TT * m];
#pragma unroll
        for (int m = 0; m < E; m++) v[m] = ccast<C>(raw[m]);
    }
    fft_reg<C, S, N, E, true>(v, s, t, a.tw);
    if (a.rtab && d.rtab >= 0) {
        const cplx<C>* tp = a.rtab + (size_t)d.rtab * K + (size_t)yc * N + t;
        cplx<C> tab[E];
#pragma unroll
        for (int m = 0; m < E; m++) tab[m] = tp[TT * m];
#pragma unroll
        for (int m = 0; m < E; m++) v[m] = cmul(v[m], tab[m]);
    }
    __syncthreads();  // the transform's LDS reads are done
#pragma unroll
    for (int m = 0; m < E; m++) s[t + TT * m] = ccast<S>(v[m]);
    __syncthreads();
    const zd* cf = a.coef + 4 * (d0.pad >> FD_PAIR_SHIFT);
    const cplx<C> rl = cmk<C>((C)cf[0].re, (C)cf[0].im), tr = cmk<C>((C)cf[1].re, (C)cf[1].im);
    const cplx<C> rr = cmk<C>((C)cf[2].re, (C)cf[2].im), tl = cmk<C>((C)cf[3].re, (C)cf[3].im);
#pragma unroll
    for (int m = 0; m < E; m++) {
        const cplx<C> own = ccast<C>(ccast<S>(v[m])), other = ccast<C>(sp[t + TT * m]);
        // k_ch_pair: Fn = c rl + tr a, Fi = a rr + tl c (a: field i, c: field i+1)
        const cplx<C> nv = side == 0 ? cadd(cmul(own, rr), cmul(tl, other)) : cadd(cmul(own, rl), cmul(tr, other));
        v[m] = ccast<C>(ccast<S>(nv));  // the stored value
    }
    if (a.xout) {  // uniform over the launch
        const uint32_t keep = aperture_bits<E, TT>(a.X, a.Y[yc], a.R2, t);
#pragma unroll
        for (int m = 0; m < E; m++)
            if (!((keep >> m) & 1u)) v[m] = cmk<C>(0, 0);
        // the transform's first barrier orders the partner's reads of s above
        // before this line's buffer is overwritten
        fft_reg<C, S, N, E, false>(v, s, t, a.tw);
    }
#pragma unroll
    for (int m = 0; m < E; m++)
        if (ok) line[TT * m] = ccast<S>(v[m]);
}

template <typename T, int N>
struct ChPairLaunch {
    // nb: entries (region pair x frequency) in a.fd
    static void run(ChPairArgs<T> a, int nb, hipStream_t s) {
        if constexpr (N >= 16) {
            using C = RegE<T, N>;
            a.rows = std::max(1, 128 / C::TT);
            const size_t smem = sizeof(cplx<typename Storage<T>::type>) * (size_t)C::LD * 2 * a.rows;
            allow_lds((const void*)k_ch_pair_rows<T, N>, smem);
            hipLaunchKernelGGL((k_ch_pair_rows<T, N>), dim3((a.yhi - a.ylo + a.rows - 1) / a.rows, nb),
                               dim3(2 * a.rows * C::TT), smem, s, a);
            BF_CHECK_LAUNCH();
        } else {
            throw Error(BF_ERR_UNSUPPORTED, "fused cheerleader rows need a power-of-two grid of at least 16");
        }
    }
};

// emission of the left-moving fields (:290-300)
template <typename T>
__global__ void k_ch_emit_left(cplx<typename Storage<T>::type>* F, const cplx<typename Storage<T>::type>* __restrict__ E,
                               int Nd, size_t K, zd tl_, zd rl_) {
    BF_TYPES(T);
    const int f = blockIdx.y;
    const cplx<C> tl = cmk<C>((C)tl_.re, (C)tl_.im), rl = cmk<C>((C)rl_.re, (C)rl_.im);
    cplx<S>* Fb = F + (size_t)f * Nd * K;
    const cplx<S>* Eb = E + (size_t)f * Nd * 2 * K;
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < K; p += (size_t)gridDim.x * blockDim.x) {
        for (int i = 0; i <= Nd - 2; i++)
            Fb[(size_t)i * K + p] = ccast<S>(cadd(ccast<C>(Fb[(size_t)i * K + p]), ccast<C>(Eb[((size_t)i * 2 + 1) * K + p])));
        const cplx<C> e = ccast<C>(Eb[((size_t)(Nd - 1) * 2 + 1) * K + p]);
        Fb[(size_t)(Nd - 2) * K + p] = ccast<S>(cadd(ccast<C>(Fb[(size_t)(Nd - 2) * K + p]), cmul(tl, e)));
        Fb[(size_t)(Nd - 1) * K + p] = ccast<S>(cadd(ccast<C>(Fb[(size_t)(Nd - 1) * K + p]), cmul(rl, e)));
    }
}
template <typename T>
__global__ void k_ch_copy_right(cplx<typename Storage<T>::type>* F, const cplx<typename Storage<T>::type>* __restrict__ E,
                                int Nd, size_t K) {
    const int f = blockIdx.y;
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < (size_t)Nd * K; p += (size_t)gridDim.x * blockDim.x) {
        const size_t i = p / K, q = p % K;
        F[(size_t)f * Nd * K + p] = E[((size_t)f * Nd + i) * 2 * K + q];
    }
}
template <typename T>
__global__ void k_amax_fields(const cplx<typename Storage<T>::type>* F, size_t n, double* amax, const int* active) {
    const int f = blockIdx.y;
    if (active && !active[f]) return;
    double m = 0.0;
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
        const zd v = ccast<double>(F[(size_t)f * n + p]);
        m = fmax(m, hypot(v.re, v.im));
    }
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0 && m > 0.0) atomic_max_pos(amax + f, m);
}

template <typename T>
void cheerleader_run(bf_ctx* ctx, const bf_setup* st, const bf_coords* co, double amin, int nmax, const double* fr,
                     int nfr, int kind, const double* emit_h, const double* reflect_h, double diskR, double* out_right,
                     double* out_left) {
    BF_TYPES(T);
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
    const size_t per = sizeof(cplx<S>) * K * Nd;
    // Gauss-Seidel is sequential over regions: a launch touches only the one or
    // two fields of a region per frequency, so batch as many frequencies as an
    // 8 GB state allows to give each launch enough work-groups.
    int fb = (int)std::max<size_t>(1, ((size_t)8 << 30) / std::max<size_t>(per * 3, 1));
    fb = std::max(1, std::min(fb, std::min(nfr, 256)));
    DevBuf F(per * fb), EM(per * 2 * fb), scr, dax, demit, dref, out(sizeof(zd) * K * fb), act(sizeof(int) * fb),
        amx(sizeof(double) * fb), dfd, dph;
    if (kind == BF_PROP_MOMENTUM) scr.alloc(per * fb);
    PropEngine<T> pe(ctx, &g, kind, diskR, surf.as<zd>());
    pe.quarter = quarter_tables_ok<T>(co, kind);
    // the pair steps' inverse row pass fused with their boundary update (register
    // path, propagator and propagatorNoTilts; option chfuse=0 runs them separately)
    bool fuse_pair = false;
    {
        // every propagation starts by masking: its masked row FFT and column pass
        // need only the aperture's row band (register path; option band=0 disables)
        auto pow2 = [](int n) { return n >= 16 && n <= 1024 && (n & (n - 1)) == 0; };
        if (opts().band && use_reg_fft() && pow2(g.nx) && pow2(g.ny) &&
            (kind == BF_PROP_PROPAGATOR || kind == BF_PROP_NOTILTS))
            aperture_band(co, diskR, pe.blo, pe.bhi);
        fuse_pair = opts().chfuse && use_reg_fft() && pow2(g.nx) &&
                    (kind == BF_PROP_PROPAGATOR || kind == BF_PROP_NOTILTS);
    }
    // The pair steps update the interior fields F[1..Nd-2] only; each is masked
    // again before it is read (its next propagation), so the steps need the band
    // rows only.  F[0] and F[Nd-1] (the outputs) come from the single steps, which
    // stay whole; the amin criterion reads every row, so it keeps them whole too.
    const bool pair_band = fuse_pair && pe.bhi > pe.blo && amin == 0.0;
    // With the band kept, the pair steps also store their fields as the next
    // propagation's masked row transform (every reader of an interior field is
    // a masked propagation).  xf[r]: F[r] holds that transform of its band rows,
    // the same for every frequency of the batch (all stay active when amin is 0);
    // option chxf=0 stores them in real space.
    const bool pair_xform = pair_band && opts().chxf;
    std::vector<char> xf(Nd, 0);
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
                fd.push_back(FieldDesc{(int64_t)(((size_t)f * Nd + i + q) * K), f * Nd + i + q, -1, f,
                                       i << FD_PAIR_SHIFT});
    }
    // boundary coefficients for the pair kernel, [i][4] = rl, tr, rr, tl
    DevBuf dcoef(sizeof(zd) * 4 * Nd);
    {
        std::vector<zd> cf(4 * Nd);
        for (int i = 0; i < Nd; i++) {
            cf[4 * i] = rl[i]; cf[4 * i + 1] = tr[i]; cf[4 * i + 2] = rr[i]; cf[4 * i + 3] = tl[i];
        }
        upload(dcoef.p, cf.data(), sizeof(zd) * cf.size(), s);
    }
    // Wavefront order (amin = 0, pair steps fused and transformed): step j of
    // iteration n (j = 0: first region, 1..Nd-3: pair (j, j+1), Nd-2: last
    // region; step j updates fields j and j+1) runs at time 2n + j.  Steps that
    // share a field keep their sequential order and the steps of one time touch
    // disjoint fields, so each time's steps share launches and the results are
    // those of the sequential sweep.  Option chwave=0 sweeps step by step.
    const bool wave = pair_xform && Nd >= 4 && opts().chwave &&
                      (size_t)(Nd / 2 + 1) * 2 * fb <= 65535;  // fields per launch within grid.y
    DevBuf dwave;
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
                               EM.as<cplx<S>>(), demit.as<zd>(), estride, estride);
        } else {
            BF_HIP(hipMemsetAsync(EM.p, 0, per * 2 * nb, s));
            if (reflect_h)
                hipLaunchKernelGGL(k_broadcast<T>, dim3(gb, nb), dim3(256), 0, s,
                                   EM.as<cplx<S>>() + ((size_t)(Nd - 1) * 2 + 1) * K, dref.as<zd>(), K, estride);
            else
                hipLaunchKernelGGL(k_dance_intro<T>, dim3(grid1d(K, 256, 512), 2 * (Nd - 1), nb), dim3(256), 0, s,
                                   EM.as<cplx<S>>() + 2 * K, Nd - 1, K, g.nx, dax.as<zd>(), g.X.as<double>(),
                                   g.Y.as<double>(), diskR * diskR, estride, (const zd*)nullptr);
        }
        BF_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_ch_copy_right<T>, dim3(grid1d(K * Nd, 256, 4096), nb), dim3(256), 0, s, F.as<cplx<S>>(),
                           EM.as<cplx<S>>(), Nd, K);
        BF_CHECK_LAUNCH();
        std::vector<int> a0(fb, 0);
        for (int f = 0; f < nb; f++) a0[f] = (1.0 >= amin) ? 1 : 0;
        const FieldDesc* D = dfd.as<FieldDesc>();
        const zd* P1 = dph.as<zd>();
        cplx<S>* Fp = F.as<cplx<S>>();
        cplx<S>* Sp = scr.as<cplx<S>>();
        auto prop_single = [&](int i, const int* active) {
            pe.propagate(Fp, D + single_off[i], nb, P1 + single_off[i], Sp, active);
        };
        auto prop_pair = [&](int i, const int* active) {
            pe.propagate(Fp, D + pair_off[i], 2 * nb, P1 + pair_off[i], Sp, active);
        };
        const uint32_t inv_rows = RF_INV | (kind == BF_PROP_PROPAGATOR ? RF_RTAB : 0u);
        // propagation of region i in the loop: a field that already holds its
        // masked row transform goes straight to the column pass
        auto prop_single_loop = [&](int i, const int* active) {
            if (!xf[i]) {
                prop_single(i, active);
                return;
            }
            pe.cols(Fp, Fp, D + single_off[i], nb, CF_FWD | CF_KMID | CF_INV, active, pe.blo, pe.bhi, true);
            pe.rows(Fp, D + single_off[i], nb, inv_rows, active);
            xf[i] = 0;
        };
        std::fill(xf.begin(), xf.end(), 0);
        // prologue (:281-300), runs for every frequency slot
        for (int i = 1; i <= Nd - 2; i++) {
            prop_single(i, nullptr);
            hipLaunchKernelGGL(k_ch_push<T>, dim3(gb, nb), dim3(256), 0, s, Fp, Nd, K, i, tr[i], rr[i], nullptr);
            BF_CHECK_LAUNCH();
        }
        hipLaunchKernelGGL(k_ch_emit_left<T>, dim3(gb, nb), dim3(256), 0, s, Fp, EM.as<cplx<S>>(), Nd, K, tl[Nd - 2],
                           rl[Nd - 2]);
        BF_CHECK_LAUNCH();
        upload(act.p, a0.data(), sizeof(int) * fb, s);
        BF_HIP(hipMemsetAsync(amx.p, 0, sizeof(double) * fb, s));
        const int* A = act.as<int>();
        // fused pair step over `entries` (region pair, frequency) entries of fdp
        auto pair_launch = [&](const FieldDesc* fdp, int entries) {
            ChPairArgs<T> pa;
            pa.F = Fp; pa.fd = fdp; pa.ny = g.ny; pa.rows = 0;
            pa.ylo = pair_band ? pe.blo : 0;
            pa.yhi = pair_band ? pe.bhi : g.ny;
            pa.tw = pe.twx;
            pa.rtab = (kind == BF_PROP_PROPAGATOR && pe.has_rtab()) ? pe.rtab.template as<cplx<C>>() : nullptr;
            pa.coef = dcoef.as<zd>();
            pa.active = A;
            pa.xout = pair_xform ? 1 : 0;
            pa.X = g.X.as<double>(); pa.Y = g.Y.as<double>(); pa.R2 = pe.R2;
            dispatch_n<ChPairLaunch, T>(g.nx, pa, entries, s);
        };
        if (wave) {
            // per time: masked row FFT of the fields in real space (r1), column pass
            // of every propagated field (c; the single regions' fields whole), inverse
            // row pass of the single regions (inv), their updates, the pair entries (p)
            struct Step {
                size_t r1, c, inv, p;
                int nr1, nc, ninv, np;
                bool first, push;
            };
            std::vector<FieldDesc> wl;
            std::vector<Step> steps;
            std::map<std::string, Step> seen;  // (step range, transformed regions) -> lists
            std::vector<char> xs(Nd, 0);
            auto add = [&](std::vector<FieldDesc>& v, int r, int pad) {
                for (int f = 0; f < nb; f++) {
                    FieldDesc d = fd[single_off[r] + f];
                    d.pad = pad;
                    v.push_back(d);
                }
            };
            for (int tau = 0; tau <= 2 * nmax + Nd - 2; tau++) {
                int jlo = std::max(0, tau - 2 * nmax), jhi = std::min(Nd - 2, tau);
                if ((tau - jlo) & 1) jlo++;
                if ((tau - jhi) & 1) jhi--;
                if (jlo > jhi) continue;
                std::string key((const char*)&jlo, sizeof(int));
                key.append((const char*)&jhi, sizeof(int));
                key.append(xs.data(), xs.size());
                auto it = seen.find(key);
                if (it != seen.end()) {
                    steps.push_back(it->second);
                    for (int j = jlo; j <= jhi; j += 2) {
                        const char v = (j == 0 || j == Nd - 2) ? 0 : 1;
                        xs[j] = xs[j + 1] = v;
                    }
                    continue;
                }
                std::vector<FieldDesc> r1, c, inv, p;
                std::vector<std::vector<FieldDesc>> pj;  // pair entries per step, ascending j
                Step st{};
                for (int j = jlo; j <= jhi; j += 2) {
                    if (j == 0 || j == Nd - 2) {
                        const int r = j == 0 ? 1 : Nd - 2;
                        if (!xs[r]) add(r1, r, 0);
                        add(c, r, FD_WHOLE);
                        add(inv, r, 0);
                        (j == 0 ? st.first : st.push) = true;
                        xs[j] = xs[j + 1] = 0;
                    } else {
                        if (!xs[j]) add(r1, j, 0);
                        if (!xs[j + 1]) add(r1, j + 1, 0);
                        add(c, j, 0);
                        add(c, j + 1, 0);
                        pj.emplace_back();
                        for (int f = 0; f < nb; f++) {
                            FieldDesc d0 = fd[single_off[j] + f], d1 = fd[single_off[j + 1] + f];
                            d0.pad = d1.pad = j << FD_PAIR_SHIFT;
                            pj.back().push_back(d0);
                            pj.back().push_back(d1);
                        }
                        xs[j] = xs[j + 1] = 1;
                    }
                }
                // The column pass takes the steps in ascending j and the pair kernel in
                // descending j: each launch then reads first what the launch before it
                // wrote last, the part of its output still in the Infinity Cache.
                for (auto it = pj.rbegin(); it != pj.rend(); ++it) p.insert(p.end(), it->begin(), it->end());
                auto put = [&](const std::vector<FieldDesc>& v, size_t& off, int& n) {
                    off = wl.size();
                    n = (int)v.size();
                    wl.insert(wl.end(), v.begin(), v.end());
                };
                put(r1, st.r1, st.nr1);
                put(c, st.c, st.nc);
                put(inv, st.inv, st.ninv);
                put(p, st.p, st.np);
                st.np /= 2;
                seen.emplace(std::move(key), st);
                steps.push_back(st);
            }
            // every descriptor addresses a field of F
            for (const FieldDesc& d : wl)
                if (d.off < 0 || (size_t)d.off + K > K * Nd * (size_t)nb)
                    throw Error(BF_ERR_ARG, "cheerleader wavefront descriptor outside the field buffer");
            dwave.alloc(sizeof(FieldDesc) * std::max<size_t>(wl.size(), 1));
            upload(dwave.p, wl.data(), sizeof(FieldDesc) * wl.size(), s);
            const FieldDesc* W = dwave.as<FieldDesc>();
            for (const Step& st : steps) {
                if (st.nr1) pe.rows(Fp, W + st.r1, st.nr1, RF_MASK | RF_FWD, A, pe.blo, pe.bhi);
                pe.cols(Fp, Fp, W + st.c, st.nc, CF_FWD | CF_KMID | CF_INV, A, pe.blo, pe.bhi, false);
                if (st.ninv) pe.rows(Fp, W + st.inv, st.ninv, inv_rows, A);
                if (st.first) {
                    hipLaunchKernelGGL(k_ch_first<T>, dim3(gb, nb), dim3(256), 0, s, Fp, Nd, K, tl[0], rl[0], A);
                    BF_CHECK_LAUNCH();
                }
                if (st.push) {
                    hipLaunchKernelGGL(k_ch_push<T>, dim3(gb, nb), dim3(256), 0, s, Fp, Nd, K, Nd - 2, tr[Nd - 2],
                                       rr[Nd - 2], A);
                    BF_CHECK_LAUNCH();
                }
                if (st.np) pair_launch(W + st.p, st.np);
            }
        }
        for (int n = 0; n <= (wave ? -1 : nmax); n++) {
            prop_single_loop(1, A);
            hipLaunchKernelGGL(k_ch_first<T>, dim3(gb, nb), dim3(256), 0, s, Fp, Nd, K, tl[0], rl[0], A);
            BF_CHECK_LAUNCH();
            xf[0] = xf[1] = 0;
            for (int i = 1; i <= Nd - 3; i++) {
                if (fuse_pair) {
                    if (pair_xform) {
                        // masked row FFT of the fields still in real space, then the column pass
                        const uint32_t fwd = RF_MASK | RF_FWD;
                        if (!xf[i] && !xf[i + 1])
                            pe.rows(Fp, D + pair_off[i], 2 * nb, fwd, A, pe.blo, pe.bhi);
                        else if (!xf[i])
                            pe.rows(Fp, D + single_off[i], nb, fwd, A, pe.blo, pe.bhi);
                        else if (!xf[i + 1])
                            pe.rows(Fp, D + single_off[i + 1], nb, fwd, A, pe.blo, pe.bhi);
                        pe.cols(Fp, Fp, D + pair_off[i], 2 * nb, CF_FWD | CF_KMID | CF_INV, A, pe.blo, pe.bhi, false);
                    } else {
                        pe.propagate_head(Fp, D + pair_off[i], 2 * nb, A, pair_band);
                    }
                    pair_launch(D + pair_off[i], nb);
                    xf[i] = xf[i + 1] = pair_xform ? 1 : 0;
                } else {
                    prop_pair(i, A);
                    hipLaunchKernelGGL(k_ch_pair<T>, dim3(gb, nb), dim3(256), 0, s, Fp, Nd, K, i, rl[i], tr[i], rr[i],
                                       tl[i], A);
                    BF_CHECK_LAUNCH();
                }
            }
            prop_single_loop(Nd - 2, A);
            hipLaunchKernelGGL(k_ch_push<T>, dim3(gb, nb), dim3(256), 0, s, Fp, Nd, K, Nd - 2, tr[Nd - 2], rr[Nd - 2],
                               A);
            BF_CHECK_LAUNCH();
            xf[Nd - 2] = xf[Nd - 1] = 0;
            if (amin != 0.0) {
                hipLaunchKernelGGL(k_amax_fields<T>, dim3(grid1d(K * Nd, 256, 1024), nb), dim3(256), 0, s, Fp,
                                   K * Nd, amx.as<double>(), A);
                launch_amin_check(s, act.as<int>(), amx.as<double>(), amin, nb);
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

void resolve_timers(bf_ctx* ctx) {
    for (auto& kv : ctx->timers) {
        bf_kernel_timer& t = kv.second;
        for (auto& pr : t.pending) {
            BF_HIP(hipEventSynchronize(pr.second));
            float m = 0.f;
            BF_HIP(hipEventElapsedTime(&m, pr.first, pr.second));
            t.ms += m;
            (void)hipEventDestroy(pr.first);
            (void)hipEventDestroy(pr.second);
        }
        t.pending.clear();
    }
}

void clear_timers(bf_ctx* ctx) {
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& kv : ctx->timers)
        for (auto& pr : kv.second.pending) {
            (void)hipEventDestroy(pr.first);
            (void)hipEventDestroy(pr.second);
        }
    ctx->timers.clear();
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
template void cheerleader_run<f32pure>(bf_ctx*, const bf_setup*, const bf_coords*, double, int, const double*, int,
                                       int, const double*, const double*, double, double*, double*);
template void cheerleader_run<float>(bf_ctx*, const bf_setup*, const bf_coords*, double, int, const double*, int, int,
                                     const double*, const double*, double, double*, double*);

}  // namespace bf
