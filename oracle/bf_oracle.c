/*
 * bf_oracle.c -- FP64 CPU restatement of BoostFractor.jl's propagation solvers.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: it is compiled into
 * oracle/libbf_oracle.so and loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, never by the product library.  The product path
 * (boostfractor.jl_amd/) never links, imports or executes anything here.
 *
 * It follows the reference operation for operation (same op order, same
 * signed-zero handling, same mask inequalities, same loop structure), with
 * these third-party pieces restated from their published algorithms:
 *   - FFTW fft!/ifft! (src/boost_fractor.jl:163,177): unnormalised forward DFT
 *     (exp(-2 pi i jk/N)), backward DFT scaled by fl(1/(Nx*Ny)).  Radix-2
 *     iterative for N = 2^k, direct O(N^2) DFT for any other N.  Twiddles are
 *     the table w[m] = exp(-2 pi i m/N) evaluated by libm cos/sin.
 *   - LAPACK zgetrf/zgetrs/zgetri (src/boost_fractor_transformer.jl:247,346,356):
 *     LU with partial pivoting, pivot chosen by max(|re|+|im|) as izamax does.
 *   - SpecialFunctions.besselj / FunctionZeros.besselj_zero
 *     (src/boost_fractor_transformer.jl:91,94): glibc jn(), zeros bracketed by
 *     a sign scan (zero spacing of J_n exceeds pi for n >= 1) then bisection.
 *   - Julia Complex{Float64} sqrt / exp / * semantics, including signed zeros
 *     (trap: real eps is stored as re - 0.0im, src/boost_fractor.jl:98,140).
 *
 * What pins each function (tests/test_oracle_goldens.py, tests/test_oracle_pins.py;
 * goldens transcribed from the reference's test/runtests.jl and notebooks):
 *   bfo_wavelength             closed form c/f
 *   bfo_seed_reflectivities    goldens runtests.jl:211-217 (r from eps), Fresnel form in the pins
 *   bfo_propagate              propagator / NoTilts: goldens :211, :215; 1D: golden :214 and its
 *                              closed-form phase; tilt and surface: closed form on a non-square
 *                              grid; MomentumSpace: identities against numpy's FFT and NoTilts
 *   bfo_dance_intro(_ex)       goldens (every dancer golden starts here); B field and velocity:
 *                              in-place write-back and InexactError cases
 *   bfo_dancer(_sweep)         goldens :214, :215; reflect / emit linearity, history and reverse
 *                              sum, amin stop against brute-force 1D bounce matrices, the
 *                              effective-reflection recursion, convergence to the cheerleader
 *   bfo_cheerleader            goldens :211, :212; reflect / left output against the
 *                              effective-reflection recursion, energy conservation
 *   bfo_besselj_zero           tabulated zeros (Abramowitz & Stegun 9.5)
 *   bfo_seed_modes             goldens :168, Beam Shapes M7 / M4
 *   bfo_field2modes            goldens :168, Beam Shapes; velocity phase closed form
 *   bfo_modes2field            golden (modes2field of the :168 coefficients)
 *   bfo_propagation_matrix     golden :171; through the transformer goldens
 *   bfo_transformer(_sweep)    golden :217; refl against the effective-reflection recursion
 *   bfo_transformer_trace_back golden :221 (Deps = 9 geometry)
 *   bfo_gauss_profile          golden runtests.jl:47
 *   bfo_unnormalized_match     closed forms (norm, conjugate symmetry, sum)
 * No function is left unpinned.
 *
 * Layout conventions (identical to the C-ABI in include/boostfractor_hip.h):
 * complex = interleaved (re, im) doubles; a field is [ny][nx] with x fastest,
 * i.e. the Julia column-major E[x, y]; stacked fields put the field index
 * slowest.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double re, im; } cx;

static inline cx C(double re, double im) { cx z; z.re = re; z.im = im; return z; }
static inline cx cadd(cx a, cx b) { return C(a.re + b.re, a.im + b.im); }
static inline cx csub(cx a, cx b) { return C(a.re - b.re, a.im - b.im); }
/* Julia: Complex * Complex, no fma contraction */
static inline cx cmul(cx a, cx b) { return C(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
/* Julia: Real * Complex, component-wise */
static inline cx rmul(double s, cx b) { return C(s * b.re, s * b.im); }
static inline cx conj_(cx a) { return C(a.re, -a.im); }
static inline double cabs2(cx a) { return a.re * a.re + a.im * a.im; }
/* Julia: (-1im) * z with -1im promoted to Complex{Float64}(0,-1) */
static inline cx neg_i(cx z) { return cmul(C(0.0, -1.0), z); }

/* Smith's algorithm; Julia uses a robust variant that agrees to ~1 ulp */
static cx cdiv(cx a, cx b) {
    if (fabs(b.re) >= fabs(b.im)) {
        double r = b.im / b.re, d = b.re + b.im * r;
        return C((a.re + a.im * r) / d, (a.im - a.re * r) / d);
    } else {
        double r = b.re / b.im, d = b.re * r + b.im;
        return C((a.re * r + a.im) / d, (a.im * r - a.re) / d);
    }
}

/* Julia Base sqrt(::Complex{Float64}) (principal branch, sign of imag follows
 * the sign of the (possibly zero) imaginary input) */
static cx csqrt_j(cx z) {
    double x = z.re, y = z.im;
    if (x == 0.0 && y == 0.0) return C(0.0, y);
    if (isinf(y)) return C(INFINITY, y);
    if (isnan(x) || isnan(y)) return C(NAN, NAN);
    double rho = sqrt((hypot(x, y) + fabs(x)) * 0.5);
    double xi = rho, eta = y;
    if (rho != 0.0) {
        if (isfinite(eta)) eta = (eta / rho) * 0.5;
        if (x < 0.0) {
            xi = fabs(eta);
            eta = copysign(rho, y);
        }
    }
    return C(xi, eta);
}

/* Julia Base exp(::Complex{Float64}) */
static cx cexp_j(cx z) {
    double er = exp(z.re);
    if (z.im == 0.0) return C(er, z.im);
    return C(er * cos(z.im), er * sin(z.im));
}

static inline cx cload(const double *p, size_t i) { return C(p[2 * i], p[2 * i + 1]); }
static inline void cstore(double *p, size_t i, cx v) { p[2 * i] = v.re; p[2 * i + 1] = v.im; }

/* pairwise summation like Julia's sum() over arrays (block 1024) */
static cx psum(const cx *v, size_t n) {
    if (n <= 1024) {
        cx s = C(0.0, 0.0);
        for (size_t i = 0; i < n; i++) s = cadd(s, v[i]);
        return s;
    }
    size_t h = n / 2;
    return cadd(psum(v, h), psum(v + h, n - h));
}

/* ----------------------------------------------------------------------- */
/* FFT (FFTW restated)                                                      */
/* ----------------------------------------------------------------------- */
typedef struct {
    int n;
    int pow2;
    cx *tw;  /* w[m] = exp(-2 pi i m / n) */
    cx *mat; /* non power of two: DFT matrix mat[k*n + j] = w[(j*k) mod n] */
} fft1;

static void fft1_init(fft1 *p, int n) {
    p->n = n;
    p->pow2 = (n & (n - 1)) == 0;
    p->tw = (cx *)malloc(sizeof(cx) * (size_t)(n > 0 ? n : 1));
    for (int m = 0; m < n; m++) {
        double a = 2.0 * M_PI * (double)m / (double)n;
        p->tw[m] = C(cos(a), -sin(a));
    }
    p->mat = NULL;
    if (!p->pow2 && n > 1) {
        p->mat = (cx *)malloc(sizeof(cx) * (size_t)n * n);
        for (int k = 0; k < n; k++)
            for (int j = 0; j < n; j++) p->mat[(size_t)k * n + j] = p->tw[(int)(((int64_t)j * k) % n)];
    }
}
static void fft1_free(fft1 *p) { free(p->tw); free(p->mat); }

/* in-place 1D transform of n elements at data[0], data[s], ...; scratch >= n */
static void fft1_run(const fft1 *p, cx *data, size_t s, int inverse, cx *buf) {
    int n = p->n;
    if (n <= 1) return;
    for (int j = 0; j < n; j++) buf[j] = data[(size_t)j * s];
    if (p->pow2) {
        int lg = 0;
        while ((1 << lg) < n) lg++;
        for (int j = 0; j < n; j++) { /* bit reversal */
            int r = 0;
            for (int b = 0; b < lg; b++) r |= ((j >> b) & 1) << (lg - 1 - b);
            if (r > j) { cx t = buf[j]; buf[j] = buf[r]; buf[r] = t; }
        }
        for (int len = 2; len <= n; len <<= 1) {
            int half = len / 2, step = n / len;
            for (int st = 0; st < n; st += len) {
                for (int k = 0; k < half; k++) {
                    cx w = p->tw[k * step];
                    if (inverse) w = conj_(w);
                    cx u = buf[st + k];
                    cx v = cmul(buf[st + k + half], w);
                    buf[st + k] = cadd(u, v);
                    buf[st + k + half] = csub(u, v);
                }
            }
        }
        for (int j = 0; j < n; j++) data[(size_t)j * s] = buf[j];
    } else {
        cx *out = buf + n;
        for (int k = 0; k < n; k++) {
            const cx *row = p->mat + (size_t)k * n;
            cx acc = C(0.0, 0.0);
            if (inverse)
                for (int j = 0; j < n; j++) acc = cadd(acc, cmul(buf[j], conj_(row[j])));
            else
                for (int j = 0; j < n; j++) acc = cadd(acc, cmul(buf[j], row[j]));
            out[k] = acc;
        }
        for (int j = 0; j < n; j++) data[(size_t)j * s] = out[j];
    }
}

typedef struct {
    int nx, ny;
    fft1 fx, fy;
} fft2;

static void fft2_init(fft2 *p, int nx, int ny) {
    p->nx = nx; p->ny = ny;
    fft1_init(&p->fx, nx);
    fft1_init(&p->fy, ny);
}
static void fft2_free(fft2 *p) { fft1_free(&p->fx); fft1_free(&p->fy); }

/* fft! (inverse=0) or ifft! (inverse=1, then scaled by fl(1/(nx*ny))) */
static void fft2_run(const fft2 *p, cx *E, int inverse) {
    int nx = p->nx, ny = p->ny;
    int nmax = nx > ny ? nx : ny;
    cx *buf = (cx *)malloc(sizeof(cx) * 2 * (size_t)nmax);
    for (int y = 0; y < ny; y++) fft1_run(&p->fx, E + (size_t)y * nx, 1, inverse, buf);
    for (int x = 0; x < nx; x++) fft1_run(&p->fy, E + x, (size_t)nx, inverse, buf);
    free(buf);
    if (inverse) {
        double sc = 1.0 / (double)((int64_t)nx * ny);
        size_t K = (size_t)nx * ny;
        for (size_t i = 0; i < K; i++) E[i] = rmul(sc, E[i]);
    }
}

/* ----------------------------------------------------------------------- */
/* Coordinates and propagators (src/boost_fractor.jl)                       */
/* ----------------------------------------------------------------------- */
typedef struct {
    int nx, ny;
    const double *X, *Y, *kX, *kY;
} coords_t;

/* wavelength (src/boost_fractor.jl:16-19) */
double bfo_wavelength(double f) { return 299792458.0 / f; }

/* 2*pi/lambda*sqrt(eps) evaluated left to right as Julia does */
static cx k0_of(cx eps, double lambda) { return rmul((2.0 * M_PI) / lambda, csqrt_j(eps)); }

/* propagatorNoTilts (src/boost_fractor.jl:157-179) */
static void prop_notilts(cx *E, const coords_t *c, const fft2 *pl, double dz, double diskR, cx eps,
                         double lambda) {
    int nx = c->nx, ny = c->ny;
    double R2 = diskR * diskR;
    for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
            double xx = c->X[x], yy = c->Y[y];
            if (!(fabs(xx * xx + yy * yy) < R2)) E[(size_t)y * nx + x] = C(0.0, 0.0);
        }
    fft2_run(pl, E, 0);
    cx k0 = k0_of(eps, lambda);
    cx k0sq = cmul(k0, k0);
    int hx = nx / 2, hy = ny / 2; /* fftshift by floor(N/2) */
    for (int my = 0; my < ny; my++) {
        double Ky = c->kY[(my + hy) % ny];
        for (int mx = 0; mx < nx; mx++) {
            double Kx = c->kX[(mx + hx) % nx];
            cx arg = C(k0sq.re - Kx * Kx, k0sq.im);
            arg = C(arg.re - Ky * Ky, arg.im);
            cx kz = csqrt_j(arg);
            cx ph = cexp_j(rmul(dz, neg_i(kz)));
            size_t i = (size_t)my * nx + mx;
            E[i] = cmul(E[i], ph);
        }
    }
    fft2_run(pl, E, 1);
}

static void apply_tilt(cx *E, const coords_t *c, cx k0, double tx, double ty) {
    cx a = neg_i(k0);
    for (int y = 0; y < c->ny; y++)
        for (int x = 0; x < c->nx; x++) {
            cx px = cexp_j(rmul(c->X[x], rmul(tx, a)));
            cx py = cexp_j(rmul(c->Y[y], rmul(ty, a)));
            size_t i = (size_t)y * c->nx + x;
            E[i] = cmul(E[i], cmul(px, py));
        }
}

static void apply_surface(cx *E, const coords_t *c, cx k0, const double *surface) {
    cx a = neg_i(k0);
    size_t K = (size_t)c->nx * c->ny;
    for (size_t i = 0; i < K; i++) {
        cx s = surface ? cload(surface, i) : C(0.0, 0.0);
        E[i] = cmul(E[i], cexp_j(cmul(a, s)));
    }
}

enum { PROP_PROPAGATOR = 0, PROP_NOTILTS = 1, PROP_1D = 2, PROP_MOMENTUM = 3 };

/* propagator (src/boost_fractor.jl:137-149) */
static void prop_full(cx *E, const coords_t *c, const fft2 *pl, double dz, double diskR, cx eps, double tx,
                      double ty, const double *surface, double lambda) {
    if (eps.im == 0.0) eps.im = -0.0;
    cx k0 = k0_of(eps, lambda);
    prop_notilts(E, c, pl, dz, diskR, eps, lambda);
    apply_tilt(E, c, k0, tx, ty);
    apply_surface(E, c, k0, surface);
}

/* propagatorMomentumSpace (src/boost_fractor.jl:187-210) */
static void prop_momentum(cx *E, const coords_t *c, const fft2 *pl, double dz, double diskR, cx eps,
                          double tx, double ty, double lambda) {
    int nx = c->nx, ny = c->ny;
    size_t K = (size_t)nx * ny;
    cx k0 = k0_of(eps, lambda);
    cx k0sq = cmul(k0, k0);
    for (int jy = 0; jy < ny; jy++)
        for (int jx = 0; jx < nx; jx++) {
            double Kx = c->kX[jx], Ky = c->kY[jy];
            cx arg = C(k0sq.re - Kx * Kx, k0sq.im);
            arg = C(arg.re - Ky * Ky, arg.im);
            cx ph = cexp_j(rmul(dz, neg_i(csqrt_j(arg))));
            size_t i = (size_t)jy * nx + jx;
            E[i] = cmul(E[i], ph);
        }
    /* ifftshift: unshifted[m] = shifted[(m + floor(N/2)) mod N] */
    cx *T = (cx *)malloc(sizeof(cx) * K);
    int hx = nx / 2, hy = ny / 2;
    for (int my = 0; my < ny; my++)
        for (int mx = 0; mx < nx; mx++)
            T[(size_t)my * nx + mx] = E[(size_t)((my + hy) % ny) * nx + (mx + hx) % nx];
    fft2_run(pl, T, 1);
    double R2 = diskR * diskR;
    for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
            double xx = c->X[x], yy = c->Y[y];
            if (!(fabs(xx * xx + yy * yy) < R2)) T[(size_t)y * nx + x] = C(0.0, 0.0);
        }
    if (tx != 0.0 || ty != 0.0) apply_tilt(T, c, k0, tx, ty);
    fft2_run(pl, T, 0);
    /* fftshift: shifted[j] = unshifted[(j - floor(N/2)) mod N] */
    for (int jy = 0; jy < ny; jy++)
        for (int jx = 0; jx < nx; jx++)
            E[(size_t)jy * nx + jx] = T[(size_t)((jy - hy + ny) % ny) * nx + (jx - hx + nx) % nx];
    free(T);
}

/* propagator1D (src/boost_fractor.jl:218-227) */
static void prop_1d(cx *E, const coords_t *c, double dz, cx eps, double lambda) {
    cx k0 = k0_of(eps, lambda);
    cx ph = cexp_j(rmul(dz, neg_i(k0)));
    size_t K = (size_t)c->nx * c->ny;
    for (size_t i = 0; i < K; i++) E[i] = cmul(E[i], ph);
}

static void prop_dispatch(int kind, cx *E, const coords_t *c, const fft2 *pl, double dz, double diskR, cx eps,
                          double tx, double ty, const double *surface, double lambda) {
    switch (kind) {
    case PROP_PROPAGATOR: prop_full(E, c, pl, dz, diskR, eps, tx, ty, surface, lambda); break;
    case PROP_NOTILTS: prop_notilts(E, c, pl, dz, diskR, eps, lambda); break;
    case PROP_1D: prop_1d(E, c, dz, eps, lambda); break;
    default: prop_momentum(E, c, pl, dz, diskR, eps, tx, ty, lambda); break;
    }
}

/* Exported single propagation: E is complex [ny][nx], overwritten with the result. */
int bfo_propagate(int kind, double *E, int nx, int ny, const double *X, const double *Y, const double *kX,
                  const double *kY, double dz, double diskR, double eps_re, double eps_im, double tx, double ty,
                  const double *surface, double lambda) {
    coords_t c = {nx, ny, X, Y, kX, kY};
    fft2 pl;
    fft2_init(&pl, nx, ny);
    prop_dispatch(kind, (cx *)E, &c, &pl, dz, diskR, C(eps_re, eps_im), tx, ty, surface, lambda);
    fft2_free(&pl);
    return 0;
}

/* ----------------------------------------------------------------------- */
/* Setup                                                                    */
/* ----------------------------------------------------------------------- */
typedef struct {
    int nreg;               /* length(distance) */
    const double *distance; /* [nreg] */
    const double *r;        /* complex [nreg-1] */
    const double *eps;      /* complex [nreg] */
    const double *tilt_x, *tilt_y;
    const double *surfaces; /* complex [nreg][ny][nx] or NULL (flat) */
} setup_t;

/* SeedSetupBoundaries reflectivities (src/boost_fractor.jl:97-102): eps is
 * modified in place to re - 0.0im when its imaginary part is zero, then
 * r = [1, (sqrt(e[i-1]) - sqrt(e[i])) / (sqrt(e[i-1]) + sqrt(e[i])) for i = 3..n] */
int bfo_seed_reflectivities(int nreg, double *eps, double *r_out) {
    for (int i = 0; i < nreg; i++)
        if (eps[2 * i + 1] == 0.0) eps[2 * i + 1] = -0.0;
    cstore(r_out, 0, C(1.0, 0.0));
    for (int i = 2; i < nreg; i++) {
        cx a = csqrt_j(cload(eps, i - 1)), b = csqrt_j(cload(eps, i));
        cstore(r_out, i - 1, cdiv(csub(a, b), cadd(a, b)));
    }
    return 0;
}

/* dance_intro (src/boost_fractor_recurprop.jl:153-201), default bfield, velocity 0.
 * nd = length(bdry.distance); eps complex [nd]; r1 = bdry.r[1].
 * out: complex [nd][2][ny][nx] */
static void dance_intro_c(int nd, const cx *eps, cx r1, const coords_t *c, double diskR, cx *out) {
    int nx = c->nx, ny = c->ny;
    size_t K = (size_t)nx * ny;
    double R2 = diskR * diskR;
    for (int n = 0; n < nd; n++) {
        cx axr;
        if (n == 0) {
            axr = (r1.re == 1.0 && r1.im == 0.0) ? C(-1.0, 0.0) : C(0.0, 0.0);
        } else {
            cx ei = eps[n - 1], em = eps[n];
            cx den = cadd(cmul(ei, csqrt_j(em)), cmul(em, csqrt_j(ei)));
            axr = cdiv(cmul(csqrt_j(em), csub(C(1.0, 0.0), cdiv(ei, em))), den);
            axr = C(axr.re, axr.im + 0.0);
        }
        cx ei = eps[n];
        cx em = (n == nd - 1) ? C(1.0, 0.0) : eps[n + 1];
        cx den = cadd(cmul(ei, csqrt_j(em)), cmul(em, csqrt_j(ei)));
        cx axl = cdiv(cmul(csqrt_j(ei), csub(C(1.0, 0.0), cdiv(em, ei))), den);
        axl = C(axl.re, axl.im + 0.0);
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                double xx = c->X[x], yy = c->Y[y];
                double b = (xx * xx + yy * yy > R2) ? 0.0 : 1.0;
                size_t i = (size_t)y * nx + x;
                out[(size_t)(2 * n) * K + i] = rmul(b, axr);
                out[(size_t)(2 * n + 1) * K + i] = rmul(b, axl);
            }
    }
}

int bfo_dance_intro(int nd, const double *eps, double r1_re, double r1_im, int nx, int ny, const double *X,
                    const double *Y, double diskR, double *out) {
    coords_t c = {nx, ny, X, Y, NULL, NULL};
    dance_intro_c(nd, (const cx *)eps, C(r1_re, r1_im), &c, diskR, (cx *)out);
    return 0;
}

/* dance_intro with a user B field and/or an axion velocity
 * (src/boost_fractor_recurprop.jl:153-201, options :157-166).
 * bfield: complex [nd+1][ny][nx], or NULL for the default real mask x^2+y^2 > R^2 ? 0 : 1.
 * velocity_x != 0 multiplies the B field by exp(-1im*Ma*(-velocity_x)*x), Ma = 2pi/lambda
 * (:162-166); with the default Float64 mask the reference's in-place `bfield .*=` throws
 * InexactError: returns -1 then.  bfield_out (optional): the multiplied B field, which the
 * reference leaves in the caller's array. */
int bfo_dance_intro_ex(int nd, const double *eps, double r1_re, double r1_im, int nx, int ny, const double *X,
                       const double *Y, double diskR, const double *bfield, double velocity_x, double f,
                       double *bfield_out, double *out) {
    size_t K = (size_t)nx * ny;
    if (bfield == NULL) {
        if (velocity_x != 0.0) return -1;
        return bfo_dance_intro(nd, eps, r1_re, r1_im, nx, ny, X, Y, diskR, out);
    }
    cx *b = (cx *)malloc(sizeof(cx) * (size_t)(nd + 1) * K);
    memcpy(b, bfield, sizeof(cx) * (size_t)(nd + 1) * K);
    if (velocity_x != 0.0) {
        double lambda = bfo_wavelength(f);
        double ma = 2.0 * M_PI / lambda;
        /* -1im*Ma*(-velocity_x)*x, evaluated left to right in complex arithmetic */
        cx a = cmul(C(0.0, -1.0), C(ma, 0.0));
        a = C(a.re * (-velocity_x), a.im * (-velocity_x));
        for (int z = 0; z <= nd; z++)
            for (int y = 0; y < ny; y++)
                for (int x = 0; x < nx; x++) {
                    cx ph = cexp_j(C(a.re * X[x], a.im * X[x]));
                    size_t i = (size_t)z * K + (size_t)y * nx + x;
                    b[i] = cmul(b[i], ph);
                }
    }
    if (bfield_out) memcpy(bfield_out, b, sizeof(cx) * (size_t)(nd + 1) * K);
    /* amplitudes as dance_intro_c, then ax (+0.0im) .* bfield[n] / bfield[n+1] */
    cx *ones_ = (cx *)malloc(sizeof(cx) * (size_t)nd * 2 * K);
    coords_t c = {nx, ny, X, Y, NULL, NULL};
    dance_intro_c(nd, (const cx *)eps, C(r1_re, r1_im), &c, INFINITY, ones_);  /* mask of all ones */
    cx *o = (cx *)out;
    for (int n = 0; n < nd; n++)
        for (size_t i = 0; i < K; i++) {
            o[(size_t)(2 * n) * K + i] = cmul(ones_[(size_t)(2 * n) * K + i], b[(size_t)n * K + i]);
            o[(size_t)(2 * n + 1) * K + i] = cmul(ones_[(size_t)(2 * n + 1) * K + i], b[(size_t)(n + 1) * K + i]);
        }
    free(ones_);
    free(b);
    return 0;
}

/* ----------------------------------------------------------------------- */
/* antenna helpers (src/beam_gaussian.jl:24-37, src/beam_coupling.jl:13-15) */
/* ----------------------------------------------------------------------- */
/* gauss_profile: omega0/omega(z) .* exp.(-radius.^2/omega(z)^2)
 *                .* exp.(-1im .* (k*z .+ k*radius.^2. /(2*R.(z)) .- Psi.(z)))
 * out: complex [ny][nx]. */
int bfo_gauss_profile(int nx, int ny, const double *X, const double *Y, double z, double omega0, double f,
                      double *out) {
    double lambda = bfo_wavelength(f);
    double k = 2.0 * M_PI / lambda;
    double zR = M_PI * omega0 * omega0 / lambda;
    double om = omega0 * sqrt(1.0 + (z / zR) * (z / zR));
    double Rz = z * (1.0 + (zR / z) * (zR / z));
    double psi = atan(z / zR);
    double amp0 = omega0 / om;
    double om2 = om * om;
    cx *o = (cx *)out;
    for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
            double r = sqrt(X[x] * X[x] + Y[y] * Y[y]);
            double r2 = r * r;
            double ph = k * z + k * r2 / (2.0 * Rz) - psi;
            /* -1im .* ph: Complex{Int}(0,-1) * ph = (0*ph, -ph) */
            cx e = cexp_j(C(0.0 * ph, -ph));
            double a = amp0 * exp(-r2 / om2);
            o[(size_t)y * nx + x] = rmul(a, e);
        }
    return 0;
}

/* unnormalized_match: sum(Eout .* conj(Eantenna)) .* dx .* dy, Julia's pairwise sum
 * (blocks of 1024 summed sequentially, halves combined recursively). */
static cx pairwise_dot(const cx *a, const cx *b, size_t lo, size_t hi) {
    if (hi - lo <= 1024) {
        cx s = C(0.0, 0.0);
        for (size_t i = lo; i < hi; i++) s = cadd(s, cmul(a[i], conj_(b[i])));
        return s;
    }
    size_t mid = lo + (hi - lo + 1) / 2; /* Julia: mid = (first + last) >>> 1, left half inclusive */
    return cadd(pairwise_dot(a, b, lo, mid), pairwise_dot(a, b, mid, hi));
}

int bfo_unnormalized_match(int64_t n, const double *Eout, const double *Eant, double dx, double dy, double *out) {
    cx s = n > 0 ? pairwise_dot((const cx *)Eout, (const cx *)Eant, 0, (size_t)n) : C(0.0, 0.0);
    s = C(s.re * dx, s.im * dx);
    s = C(s.re * dy, s.im * dy);
    out[0] = s.re;
    out[1] = s.im;
    return 0;
}

/* ----------------------------------------------------------------------- */
/* dancer (src/boost_fractor_recurprop.jl:25-146)                           */
/* ----------------------------------------------------------------------- */
/* sb is the UNtrimmed SetupBoundaries (as the reference's callers pass it).
 * emit: complex [nreg-1][2][ny][nx] or NULL; reflect: complex [ny][nx] or NULL.
 * out: returnsum -> [ny][nx]; else immediatesum -> [1][ny][nx] else [nmax+1][ny][nx].
 * Returns the number of bounces executed. */
static int dancer_c(const setup_t *sb, const coords_t *c, double amin, int nmax, double f, int kind,
                    const cx *emit, const cx *reflect, double diskR, int returnsum, int immediatesum, cx *out) {
    int nx = c->nx, ny = c->ny;
    size_t K = (size_t)nx * ny;
    int R = sb->nreg - 1; /* regions after dropping the mirror region */
    /* trimmed copies (:28-34) */
    const double *dist = sb->distance + 1;
    const cx *eps = (const cx *)sb->eps + 1;
    const double *tx = sb->tilt_x + 1, *ty = sb->tilt_y + 1;
    const double *surf = sb->surfaces ? sb->surfaces + 2 * K : NULL;
    cx *r = (cx *)malloc(sizeof(cx) * (size_t)(R + 1));
    for (int i = 0; i < R; i++) r[i] = cload(sb->r, i);
    r[R] = C(0.0, 0.0); /* append!(bdry.r, 0.0) */

    cx *fields = (cx *)calloc((size_t)R * 2 * K, sizeof(cx));
    int nslices = immediatesum ? 1 : nmax + 1;
    cx *Eout = (cx *)calloc((size_t)nslices * K, sizeof(cx));
    if (reflect == NULL) {
        if (emit == NULL)
            dance_intro_c(R, eps, r[0], c, diskR, fields);
        else
            memcpy(fields, emit, sizeof(cx) * (size_t)R * 2 * K);
    } else {
        memcpy(fields + (size_t)(2 * (R - 1) + 1) * K, reflect, sizeof(cx) * K);
    }
    double lambda = bfo_wavelength(f);
    fft2 pl;
    fft2_init(&pl, nx, ny);
    cx *FAL = (cx *)malloc(sizeof(cx) * K);
    int rm = 0, lm = 1; /* rightmoving = 1, leftmoving = 2 (0-based slots) */
    int n = 0;
    double a = 1.0;
#define FLD(reg, slot) (fields + ((size_t)(reg) * 2 + (slot)) * K)
    while (n <= nmax && a >= amin) {
        for (int i = 1; i <= R + 1; i++) { /* boundaries, 1-based as in the reference */
            if (i > 1) {
                int j = i - 2; /* region i-1, 0-based */
                prop_dispatch(kind, FLD(j, rm), c, &pl, dist[j], diskR, eps[j], tx[j], ty[j],
                              surf ? surf + 2 * (size_t)j * K : NULL, lambda);
            }
            if (i < R + 1) {
                int j = i - 1; /* region i */
                prop_dispatch(kind, FLD(j, lm), c, &pl, dist[j], diskR, eps[j], tx[j], ty[j],
                              surf ? surf + 2 * (size_t)j * K : NULL, lambda);
            }
            cx ri = r[i - 1];
            if (i == 1) {
                cx m = C(-ri.re, -ri.im);
                cx *L = FLD(0, lm);
                for (size_t p = 0; p < K; p++) L[p] = cmul(L[p], m);
            } else if (i == R + 1) {
                cx *A = FLD(i - 2, rm);
                cx t = C(1.0 + ri.re, ri.im);
                cx *dst = Eout + (immediatesum ? 0 : (size_t)n * K);
                for (size_t p = 0; p < K; p++) {
                    if (immediatesum)
                        dst[p] = cadd(dst[p], cmul(t, A[p]));
                    else
                        dst[p] = cmul(t, A[p]);
                    A[p] = cmul(A[p], ri);
                }
            } else {
                cx *A = FLD(i - 2, rm), *B = FLD(i - 1, lm);
                cx omr = C(1.0 - ri.re, -ri.im), opr = C(1.0 + ri.re, ri.im), mr = C(-ri.re, -ri.im);
                memcpy(FAL, A, sizeof(cx) * K);
                for (size_t p = 0; p < K; p++) {
                    A[p] = cmul(A[p], ri);
                    A[p] = cadd(A[p], cmul(omr, B[p]));
                    B[p] = cmul(B[p], mr);
                    B[p] = cadd(B[p], cmul(opr, FAL[p]));
                }
            }
        }
        int t = lm; lm = rm; rm = t;
        if (amin != 0.0) {
            a = 0.0;
            for (size_t p = 0; p < (size_t)R * 2 * K; p++) {
                double v = hypot(fields[p].re, fields[p].im);
                if (v > a || isnan(v)) a = v;
            }
        }
        n += 1;
    }
#undef FLD
    if (returnsum) {
        /* sum(reverse(Eout, dims=3), dims=3): sequential from the last slice */
        for (size_t p = 0; p < K; p++) {
            cx s = C(0.0, 0.0);
            for (int q = nslices - 1; q >= 0; q--) s = cadd(s, Eout[(size_t)q * K + p]);
            out[p] = s;
        }
    } else {
        memcpy(out, Eout, sizeof(cx) * (size_t)nslices * K);
    }
    free(FAL); free(Eout); free(fields); free(r);
    fft2_free(&pl);
    return n;
}

static void unpack_setup(setup_t *s, int nreg, const double *distance, const double *r, const double *eps,
                         const double *tilt_x, const double *tilt_y, const double *surfaces) {
    s->nreg = nreg; s->distance = distance; s->r = r; s->eps = eps;
    s->tilt_x = tilt_x; s->tilt_y = tilt_y; s->surfaces = surfaces;
}

int bfo_dancer(int nreg, const double *distance, const double *r, const double *eps, const double *tilt_x,
               const double *tilt_y, const double *surfaces, int nx, int ny, const double *X, const double *Y,
               const double *kX, const double *kY, double amin, int nmax, double f, int kind,
               const double *emit, const double *reflect, double diskR, int returnsum, int immediatesum,
               double *out) {
    setup_t s;
    unpack_setup(&s, nreg, distance, r, eps, tilt_x, tilt_y, surfaces);
    coords_t c = {nx, ny, X, Y, kX, kY};
    return dancer_c(&s, &c, amin, nmax, f, kind, (const cx *)emit, (const cx *)reflect, diskR, returnsum,
                    immediatesum, (cx *)out);
}

/* Frequency sweep, one point per OpenMP thread (the reference's one point per
 * Distributed worker).  out: [nfreq][ny][nx] summed output. */
int bfo_dancer_sweep(int nreg, const double *distance, const double *r, const double *eps, const double *tilt_x,
                     const double *tilt_y, const double *surfaces, int nx, int ny, const double *X,
                     const double *Y, const double *kX, const double *kY, double amin, int nmax,
                     const double *freqs, int nfreq, int kind, double diskR, int nthreads, double *out) {
    setup_t s;
    unpack_setup(&s, nreg, distance, r, eps, tilt_x, tilt_y, surfaces);
    coords_t c = {nx, ny, X, Y, kX, kY};
    size_t K = (size_t)nx * ny;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
    for (int q = 0; q < nfreq; q++)
        dancer_c(&s, &c, amin, nmax, freqs[q], kind, NULL, NULL, diskR, 1, 1, (cx *)out + (size_t)q * K);
    return 0;
}

/* ----------------------------------------------------------------------- */
/* cheerleader (src/boost_fractor_recurprop.jl:226-363)                     */
/* ----------------------------------------------------------------------- */
static int cheerleader_c(const setup_t *b, const coords_t *c, double amin, int nmax, double f, int kind,
                         const cx *emit_in, const cx *reflect, double diskR, cx *out_right, cx *out_left) {
    int nx = c->nx, ny = c->ny;
    size_t K = (size_t)nx * ny;
    int Nd = b->nreg;
    double lambda = bfo_wavelength(f);
    const cx *eps = (const cx *)b->eps;
    cx *rl = (cx *)malloc(sizeof(cx) * (size_t)Nd), *rr = (cx *)malloc(sizeof(cx) * (size_t)Nd);
    cx *tl = (cx *)malloc(sizeof(cx) * (size_t)Nd), *tr = (cx *)malloc(sizeof(cx) * (size_t)Nd);
    for (int i = 0; i < Nd - 1; i++) {
        cx ri = cload(b->r, i);
        rl[i] = C(-ri.re, -ri.im);
        rr[i] = ri;
        tl[i] = C(1.0 - ri.re, -ri.im);
        tr[i] = C(1.0 + ri.re, ri.im);
    }
    cx *emit = (cx *)calloc((size_t)Nd * 2 * K, sizeof(cx));
    if (reflect == NULL) {
        if (emit_in == NULL)
            dance_intro_c(Nd - 1, eps + 1, cload(b->r, 0), c, diskR, emit + 2 * K);
        else
            memcpy(emit, emit_in, sizeof(cx) * (size_t)Nd * 2 * K);
    } else {
        memcpy(emit + (size_t)(2 * (Nd - 1) + 1) * K, reflect, sizeof(cx) * K);
    }
#define EM(reg, slot) (emit + ((size_t)(reg) * 2 + (slot)) * K)
#define FL(reg) (fields + (size_t)(reg) * K)
    cx *fields = (cx *)malloc(sizeof(cx) * (size_t)Nd * K);
    for (int i = 0; i < Nd; i++) memcpy(FL(i), EM(i, 0), sizeof(cx) * K);
    fft2 pl;
    fft2_init(&pl, nx, ny);
    const double *surf = b->surfaces;
#define PROP(j) prop_dispatch(kind, FL(j), c, &pl, b->distance[j], diskR, eps[j], b->tilt_x[j], b->tilt_y[j], \
                              surf ? surf + 2 * (size_t)(j) * K : NULL, lambda)
    /* push the right-moving emission forward (0-based regions 1..Nd-2) */
    for (int i = 1; i <= Nd - 2; i++) {
        PROP(i);
        cx *Fi = FL(i), *Fn = FL(i + 1);
        for (size_t p = 0; p < K; p++) {
            Fn[p] = cadd(Fn[p], cmul(tr[i], Fi[p]));
            Fi[p] = cmul(Fi[p], rr[i]);
        }
    }
    for (int i = 0; i <= Nd - 2; i++) {
        cx *Fi = FL(i), *Ei = EM(i, 1);
        for (size_t p = 0; p < K; p++) Fi[p] = cadd(Fi[p], Ei[p]);
    }
    {
        int i = Nd - 2;
        cx *Fi = FL(i), *Fn = FL(i + 1), *E = EM(i + 1, 1);
        for (size_t p = 0; p < K; p++) {
            Fi[p] = cadd(Fi[p], cmul(tl[i], E[p]));
            Fn[p] = cadd(Fn[p], cmul(rl[i], E[p]));
        }
    }
    cx *cpy = (cx *)malloc(sizeof(cx) * K);
    int n = 0;
    double a = 1.0;
    while (n <= nmax && a >= amin) {
        { /* first region */
            PROP(1);
            cx *F0 = FL(0), *F1 = FL(1);
            for (size_t p = 0; p < K; p++) {
                F0[p] = cadd(F0[p], cmul(tl[0], F1[p]));
                F1[p] = cmul(F1[p], rl[0]);
            }
        }
        for (int i = 1; i <= Nd - 3; i++) {
            PROP(i);
            PROP(i + 1);
            cx *Fi = FL(i), *Fn = FL(i + 1);
            memcpy(cpy, Fn, sizeof(cx) * K);
            for (size_t p = 0; p < K; p++) {
                Fn[p] = cmul(Fn[p], rl[i]);
                Fn[p] = cadd(Fn[p], cmul(tr[i], Fi[p]));
                Fi[p] = cmul(Fi[p], rr[i]);
                Fi[p] = cadd(Fi[p], cmul(tl[i], cpy[p]));
            }
        }
        {
            int i = Nd - 2;
            PROP(i);
            cx *Fi = FL(i), *Fn = FL(i + 1);
            for (size_t p = 0; p < K; p++) {
                Fn[p] = cadd(Fn[p], cmul(tr[i], Fi[p]));
                Fi[p] = cmul(Fi[p], rr[i]);
            }
        }
        if (amin != 0.0) {
            a = 0.0;
            for (size_t p = 0; p < (size_t)Nd * K; p++) {
                double v = hypot(fields[p].re, fields[p].im);
                if (v > a || isnan(v)) a = v;
            }
        }
        n += 1;
    }
#undef PROP
    memcpy(out_right, FL(Nd - 1), sizeof(cx) * K);
    if (out_left) memcpy(out_left, FL(0), sizeof(cx) * K);
#undef EM
#undef FL
    free(cpy); free(fields); free(emit); free(rl); free(rr); free(tl); free(tr);
    fft2_free(&pl);
    return n;
}

int bfo_cheerleader(int nreg, const double *distance, const double *r, const double *eps, const double *tilt_x,
                    const double *tilt_y, const double *surfaces, int nx, int ny, const double *X,
                    const double *Y, const double *kX, const double *kY, double amin, int nmax, double f,
                    int kind, const double *emit, const double *reflect, double diskR, double *out_right,
                    double *out_left) {
    setup_t s;
    unpack_setup(&s, nreg, distance, r, eps, tilt_x, tilt_y, surfaces);
    coords_t c = {nx, ny, X, Y, kX, kY};
    return cheerleader_c(&s, &c, amin, nmax, f, kind, (const cx *)emit, (const cx *)reflect, diskR,
                         (cx *)out_right, (cx *)out_left);
}

/* ----------------------------------------------------------------------- */
/* Modes (src/boost_fractor_transformer.jl:46-107, 115-167)                 */
/* ----------------------------------------------------------------------- */
/* m-th positive zero of J_nu, nu >= 0, m >= 1 (FunctionZeros.besselj_zero) */
double bfo_besselj_zero(int nu, int m) {
    double x = (nu == 0) ? 0.5 : (double)nu; /* all positive zeros lie beyond nu */
    double h = 0.25;
    double fa = jn(nu, x);
    int found = 0;
    for (;;) {
        double xb = x + h, fb = jn(nu, xb);
        if ((fa < 0.0 && fb >= 0.0) || (fa > 0.0 && fb <= 0.0)) {
            if (++found == m) {
                double lo = x, hi = xb, flo = fa;
                for (int it = 0; it < 200; it++) {
                    double mid = 0.5 * (lo + hi);
                    if (mid <= lo || mid >= hi) break;
                    double fm = jn(nu, mid);
                    if (fm == 0.0) { lo = hi = mid; break; }
                    if ((flo < 0.0) == (fm < 0.0)) { lo = mid; flo = fm; } else { hi = mid; }
                }
                return fabs(jn(nu, lo)) <= fabs(jn(nu, hi)) ? lo : hi;
            }
        }
        x = xb; fa = fb;
    }
}

/* Bessel mode patterns: patterns complex [n][ny][nx], kt complex [n],
 * n index = (m-1)(2L+1) + l + L (m = 1..M, l = -L..L) */
int bfo_seed_modes(int nx, int ny, const double *X, const double *Y, int M, int L, double diskR,
                   double *patterns, double *kt) {
    size_t K = (size_t)nx * ny;
    cx *pat = (cx *)patterns;
    cx *tmp = (cx *)malloc(sizeof(cx) * K);
    for (int m = 1; m <= M; m++)
        for (int l = -L; l <= L; l++) {
            int idx = (m - 1) * (2 * L + 1) + l + L;
            double kr = bfo_besselj_zero(abs(l), m) / diskR;
            cstore(kt, idx, C(kr, 0.0));
            cx *P = pat + (size_t)idx * K;
            for (int y = 0; y < ny; y++)
                for (int x = 0; x < nx; x++) {
                    double xx = X[x], yy = Y[y];
                    double RR = sqrt(xx * xx + yy * yy);
                    double phi = atan2(yy, xx);
                    double j = jn(l, kr * RR);
                    /* exp(-1im*l*Phi): (-1im*l) = (0, -l) times Phi, then exp */
                    cx e = cexp_j(rmul(phi, C(0.0 * l, -1.0 * l)));
                    cx v = rmul(j, e);
                    if (RR > diskR) v = rmul(0.0, v);
                    P[(size_t)y * nx + x] = v;
                }
            for (size_t i = 0; i < K; i++) tmp[i] = C(cabs2(P[i]), 0.0);
            double nrm = sqrt(psum(tmp, K).re);
            for (size_t i = 0; i < K; i++) P[i] = C(P[i].re / nrm, P[i].im / nrm);
        }
    free(tmp);
    return 0;
}

/* axion_induced_modes / field2modes (src/boost_fractor_transformer.jl:115-151,156):
 * B complex [ny][nx] or NULL (ones), velocity phase when velocity_x != 0.
 * patterns complex [n][ny][nx] (one field component). out complex [n]. */
int bfo_field2modes(int nx, int ny, const double *X, const double *Y, int n, const double *patterns,
                    const double *B, double velocity_x, double diskR, double f, double *out) {
    size_t K = (size_t)nx * ny;
    cx *b = (cx *)malloc(sizeof(cx) * K);
    cx *tmp = (cx *)malloc(sizeof(cx) * K);
    for (size_t i = 0; i < K; i++) b[i] = B ? cload(B, i) : C(1.0, 0.0);
    if (velocity_x != 0.0) {
        double Ma = 2.0 * M_PI / bfo_wavelength(f);
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                /* exp(-1im*Ma*(-v)*x) */
                cx e = cexp_j(rmul(X[x], rmul(-velocity_x, rmul(Ma, C(0.0, -1.0)))));
                size_t i = (size_t)y * nx + x;
                b[i] = cmul(b[i], e);
            }
    }
    for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
            double xx = X[x], yy = Y[y];
            size_t i = (size_t)y * nx + x;
            if (!(sqrt(xx * xx + yy * yy) <= diskR)) b[i] = rmul(0.0, b[i]);
        }
    for (size_t i = 0; i < K; i++) tmp[i] = C(cabs2(b[i]), 0.0);
    double nrm = sqrt(psum(tmp, K).re);
    for (size_t i = 0; i < K; i++) b[i] = C(b[i].re / nrm, b[i].im / nrm);
    const cx *pat = (const cx *)patterns;
    for (int k = 0; k < n; k++) {
        const cx *P = pat + (size_t)k * K;
        for (size_t i = 0; i < K; i++) tmp[i] = cmul(conj_(P[i]), b[i]);
        cstore(out, k, psum(tmp, K));
    }
    free(b); free(tmp);
    return 0;
}

/* modes2field (src/boost_fractor_transformer.jl:161-167) */
int bfo_modes2field(int K_, int n, const double *patterns, const double *coeffs, double *out) {
    size_t K = (size_t)K_;
    cx *o = (cx *)out;
    const cx *pat = (const cx *)patterns;
    for (size_t i = 0; i < K; i++) o[i] = C(0.0, 0.0);
    for (int k = 0; k < n; k++) {
        cx a = cload(coeffs, k);
        for (size_t i = 0; i < K; i++) o[i] = cadd(o[i], cmul(a, pat[(size_t)k * K + i]));
    }
    return 0;
}

/* ----------------------------------------------------------------------- */
/* Dense complex linear algebra (LAPACK restated)                           */
/* ----------------------------------------------------------------------- */
/* matrices are column-major n x n (Julia layout): A[i + n*j] */
static int lu_factor(cx *A, int n, int *piv) {
    for (int k = 0; k < n; k++) {
        int p = k;
        double best = -1.0;
        for (int i = k; i < n; i++) {
            double v = fabs(A[i + (size_t)n * k].re) + fabs(A[i + (size_t)n * k].im);
            if (v > best) { best = v; p = i; }
        }
        piv[k] = p;
        if (p != k)
            for (int j = 0; j < n; j++) {
                cx t = A[k + (size_t)n * j]; A[k + (size_t)n * j] = A[p + (size_t)n * j]; A[p + (size_t)n * j] = t;
            }
        cx d = A[k + (size_t)n * k];
        if (d.re == 0.0 && d.im == 0.0) return k + 1;
        for (int i = k + 1; i < n; i++) A[i + (size_t)n * k] = cdiv(A[i + (size_t)n * k], d);
        for (int j = k + 1; j < n; j++) {
            cx akj = A[k + (size_t)n * j];
            for (int i = k + 1; i < n; i++)
                A[i + (size_t)n * j] = csub(A[i + (size_t)n * j], cmul(A[i + (size_t)n * k], akj));
        }
    }
    return 0;
}

static void lu_solve(const cx *LU, int n, const int *piv, cx *b) {
    for (int k = 0; k < n; k++)
        if (piv[k] != k) { cx t = b[k]; b[k] = b[piv[k]]; b[piv[k]] = t; }
    for (int i = 0; i < n; i++)
        for (int k = 0; k < i; k++) b[i] = csub(b[i], cmul(LU[i + (size_t)n * k], b[k]));
    for (int i = n - 1; i >= 0; i--) {
        for (int k = i + 1; k < n; k++) b[i] = csub(b[i], cmul(LU[i + (size_t)n * k], b[k]));
        b[i] = cdiv(b[i], LU[i + (size_t)n * i]);
    }
}

static int cinv(const cx *A, int n, cx *out) {
    cx *LU = (cx *)malloc(sizeof(cx) * (size_t)n * n);
    int *piv = (int *)malloc(sizeof(int) * (size_t)n);
    memcpy(LU, A, sizeof(cx) * (size_t)n * n);
    int info = lu_factor(LU, n, piv);
    for (int j = 0; j < n; j++) {
        cx *col = out + (size_t)n * j;
        for (int i = 0; i < n; i++) col[i] = C(i == j ? 1.0 : 0.0, 0.0);
        lu_solve(LU, n, piv, col);
    }
    free(LU); free(piv);
    return info;
}

/* C = A * B, all n x n column-major */
static void cmatmul(const cx *A, const cx *B, int n, cx *Cm) {
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            cx s = C(0.0, 0.0);
            for (int k = 0; k < n; k++) s = cadd(s, cmul(A[i + (size_t)n * k], B[k + (size_t)n * j]));
            Cm[i + (size_t)n * j] = s;
        }
}

/* ----------------------------------------------------------------------- */
/* Transformer (src/boost_fractor_transformer.jl:174-421)                   */
/* ----------------------------------------------------------------------- */
typedef struct {
    int M, L, n;
    const cx *pat; /* [n][ny][nx] */
    const cx *kt;  /* [n] */
} modes_t;

/* propagation_matrix (:174-227); out column-major n x n */
static void propagation_matrix_c(double dz, double diskR, cx eps, double tx, double ty, const double *surface,
                                 double lambda, const coords_t *c, const fft2 *pl, const modes_t *md, int kind,
                                 int onlydiagonal, cx *out) {
    int n = md->n;
    size_t K = (size_t)c->nx * c->ny;
    int is_air = (eps.re == 1.0);
    cx k0 = k0_of(eps, lambda);
    cx *prop = (cx *)malloc(sizeof(cx) * K);
    cx *tmp = (cx *)malloc(sizeof(cx) * K);
    for (size_t i = 0; i < (size_t)n * n; i++) out[i] = C(0.0, 0.0);
    for (int jp = 0; jp < n; jp++) {
        memcpy(prop, md->pat + (size_t)jp * K, sizeof(cx) * K);
        if (is_air) {
            prop_dispatch(kind, prop, c, pl, dz, diskR, eps, tx, ty, surface, lambda);
        } else {
            /* efields .* tilt .* exp.(-1im*k0*surface) */
            cx a = neg_i(k0);
            for (int y = 0; y < c->ny; y++)
                for (int x = 0; x < c->nx; x++) {
                    size_t i = (size_t)y * c->nx + x;
                    cx px = cexp_j(rmul(c->X[x], rmul(tx, a)));
                    cx py = cexp_j(rmul(c->Y[y], rmul(ty, a)));
                    cx s = surface ? cload(surface, i) : C(0.0, 0.0);
                    prop[i] = cmul(cmul(prop[i], cmul(px, py)), cexp_j(cmul(a, s)));
                }
        }
        for (int j = 0; j < n; j++) {
            if (onlydiagonal && j != jp) continue;
            const cx *P = md->pat + (size_t)j * K;
            for (size_t i = 0; i < K; i++) tmp[i] = cmul(conj_(P[i]), prop[i]);
            out[j + (size_t)n * jp] = cadd(out[j + (size_t)n * jp], psum(tmp, K));
        }
    }
    if (!is_air) {
        /* diag(exp(-1im*kz*dz)) * matching, kz = sqrt(k0^2 - kt^2) */
        cx k0sq = cmul(k0, k0);
        for (int j = 0; j < n; j++) {
            cx kt2 = cmul(md->kt[j], md->kt[j]);
            cx kz = csqrt_j(csub(k0sq, kt2));
            cx d = cexp_j(rmul(dz, neg_i(kz)));
            for (int q = 0; q < n; q++) out[j + (size_t)n * q] = cmul(d, out[j + (size_t)n * q]);
        }
    }
    free(prop); free(tmp);
}

int bfo_propagation_matrix(double dz, double diskR, double eps_re, double eps_im, double tx, double ty,
                           const double *surface, double lambda, int nx, int ny, const double *X,
                           const double *Y, const double *kX, const double *kY, int n, const double *patterns,
                           const double *kt, int kind, int onlydiagonal, double *out) {
    coords_t c = {nx, ny, X, Y, kX, kY};
    modes_t md = {0, 0, n, (const cx *)patterns, (const cx *)kt};
    fft2 pl;
    fft2_init(&pl, nx, ny);
    propagation_matrix_c(dz, diskR, C(eps_re, eps_im), tx, ty, surface, lambda, &c, &pl, &md, kind, onlydiagonal,
                         (cx *)out);
    fft2_free(&pl);
    return 0;
}

/* get_boundary_matrix (:238-252): G * [P 0; 0 inv(P)], 2n x 2n column-major */
static void boundary_matrix(cx nl, cx nr, const cx *P, int n, cx *out) {
    int N2 = 2 * n;
    cx *Pi = (cx *)malloc(sizeof(cx) * (size_t)n * n);
    cinv(P, n, Pi);
    cx s = cdiv(C(1.0, 0.0), rmul(2.0, nr));
    cx a = cmul(s, cadd(nr, nl)), b = cmul(s, csub(nr, nl));
    /* dense G * blockdiag: the zero blocks contribute exact zeros */
    for (int j = 0; j < N2; j++)
        for (int i = 0; i < N2; i++) {
            int bi = i / n, ii = i % n, bj = j / n, jj = j % n;
            const cx *Mb = (bj == 0) ? P : Pi;
            cx g = (bi == bj) ? a : b;
            out[i + (size_t)N2 * j] = cmul(g, Mb[ii + (size_t)n * jj]);
        }
    free(Pi);
}

static cx eps_fix(cx e) { return (isnan(e.re) || isnan(e.im)) ? C(1e30, 0.0) : e; }

/* transformer (:285-359).  prop_mats: NULL or complex [nreg][n][n] (column-major each).
 * emit complex [n]; reflect complex [n] or NULL.  boost complex [n]; refl complex [n] (if reflect). */
static int transformer_c(const setup_t *b, const coords_t *c, const modes_t *md, double f, int kind,
                         const cx *prop_mats, double diskR, const cx *emit, const cx *reflect, cx *boost,
                         cx *refl) {
    int n = md->n, N2 = 2 * n, Nreg = b->nreg;
    size_t K = (size_t)c->nx * c->ny;
    double lambda = bfo_wavelength(f);
    cx *eps = (cx *)malloc(sizeof(cx) * (size_t)Nreg);
    for (int i = 0; i < Nreg; i++) eps[i] = eps_fix(cload(b->eps, i)); /* bdry.eps[isnan] .= 1e30 */
    cx *T = (cx *)calloc((size_t)N2 * N2, sizeof(cx));
    cx *Bm = (cx *)malloc(sizeof(cx) * (size_t)N2 * N2);
    cx *Tn = (cx *)malloc(sizeof(cx) * (size_t)N2 * N2);
    cx *P = (cx *)malloc(sizeof(cx) * (size_t)n * n);
    cx *ab = (cx *)calloc((size_t)n, sizeof(cx));
    cx *v = (cx *)malloc(sizeof(cx) * (size_t)n);
    for (int i = 0; i < N2; i++) T[i + (size_t)N2 * i] = C(1.0, 0.0);
    fft2 pl;
    fft2_init(&pl, c->nx, c->ny);
    for (int s = Nreg - 1; s >= 1; s--) {
        int ir = Nreg - s;       /* idx_reg(s) - 1, 0-based */
        int ir1 = Nreg - s - 1;  /* idx_reg(s+1) - 1 */
        /* axion_contrib(T, sqrt(eps[idx(s+1)]), sqrt(eps[idx(s)]), initial) (:258-268) */
        cx n1 = csqrt_j(eps[ir1]), n0 = csqrt_j(eps[ir]);
        cx S = rmul(0.5, csub(cdiv(C(1.0, 0.0), cmul(n1, n1)), cdiv(C(1.0, 0.0), cmul(n0, n0))));
        for (int i = 0; i < n; i++) {
            cx s21 = C(0.0, 0.0), s22 = C(0.0, 0.0);
            for (int k = 0; k < n; k++) {
                s21 = cadd(s21, cmul(T[(n + i) + (size_t)N2 * k], emit[k]));
                s22 = cadd(s22, cmul(T[(n + i) + (size_t)N2 * (n + k)], emit[k]));
            }
            ab[i] = cadd(ab[i], cmul(S, cadd(s21, s22)));
        }
        if (prop_mats)
            memcpy(P, prop_mats + (size_t)ir * n * n, sizeof(cx) * (size_t)n * n);
        else
            propagation_matrix_c(b->distance[ir], diskR, eps[ir], b->tilt_x[ir], b->tilt_y[ir],
                                 b->surfaces ? b->surfaces + 2 * (size_t)ir * K : NULL, lambda, c, &pl, md, kind, 0,
                                 P);
        boundary_matrix(n0, n1, P, n, Bm);
        cmatmul(T, Bm, N2, Tn);
        memcpy(T, Tn, sizeof(cx) * (size_t)N2 * N2);
    }
    /* boost = -T22 \ axion_beam */
    cx *T22 = (cx *)malloc(sizeof(cx) * (size_t)n * n);
    int *piv = (int *)malloc(sizeof(int) * (size_t)n);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) T22[i + (size_t)n * j] = T[(n + i) + (size_t)N2 * (n + j)];
    int info = lu_factor(T22, n, piv);
    for (int i = 0; i < n; i++) v[i] = ab[i];
    lu_solve(T22, n, piv, v);
    for (int i = 0; i < n; i++) boost[i] = C(-v[i].re, -v[i].im);
    if (reflect) {
        for (int i = 0; i < n; i++) {
            cx s21 = C(0.0, 0.0);
            for (int k = 0; k < n; k++) s21 = cadd(s21, cmul(T[(n + i) + (size_t)N2 * k], reflect[k]));
            v[i] = s21;
        }
        lu_solve(T22, n, piv, v);
        for (int i = 0; i < n; i++) refl[i] = C(-v[i].re, -v[i].im);
    }
    fft2_free(&pl);
    free(T22); free(piv); free(eps); free(T); free(Bm); free(Tn); free(P); free(ab); free(v);
    return info;
}

int bfo_transformer(int nreg, const double *distance, const double *eps, const double *tilt_x,
                    const double *tilt_y, const double *surfaces, int nx, int ny, const double *X, const double *Y,
                    const double *kX, const double *kY, int n, const double *patterns, const double *kt, double f,
                    int kind, const double *prop_mats, double diskR, const double *emit, const double *reflect,
                    double *boost, double *refl) {
    setup_t s;
    unpack_setup(&s, nreg, distance, NULL, eps, tilt_x, tilt_y, surfaces);
    coords_t c = {nx, ny, X, Y, kX, kY};
    modes_t md = {0, 0, n, (const cx *)patterns, (const cx *)kt};
    return transformer_c(&s, &c, &md, f, kind, (const cx *)prop_mats, diskR, (const cx *)emit,
                         (const cx *)reflect, (cx *)boost, (cx *)refl);
}

/* Frequency sweep of the transformer, one point per thread.  boost: [nfreq][n] */
int bfo_transformer_sweep(int nreg, const double *distance, const double *eps, const double *tilt_x,
                          const double *tilt_y, const double *surfaces, int nx, int ny, const double *X,
                          const double *Y, const double *kX, const double *kY, int n, const double *patterns,
                          const double *kt, const double *freqs, int nfreq, int kind, double diskR,
                          const double *emit, int nthreads, double *boost) {
    setup_t s;
    unpack_setup(&s, nreg, distance, NULL, eps, tilt_x, tilt_y, surfaces);
    coords_t c = {nx, ny, X, Y, kX, kY};
    modes_t md = {0, 0, n, (const cx *)patterns, (const cx *)kt};
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
    for (int q = 0; q < nfreq; q++)
        transformer_c(&s, &c, &md, freqs[q], kind, NULL, diskR, (const cx *)emit, NULL, (cx *)boost + (size_t)q * n,
                      NULL);
    return 0;
}

/* transformer_trace_back (:372-421).  fields_out complex [nreg][2][n] */
int bfo_transformer_trace_back(const double *reflected, const double *input, int nreg, const double *distance,
                               const double *eps, const double *tilt_x, const double *tilt_y,
                               const double *surfaces, int nx, int ny, const double *X, const double *Y,
                               const double *kX, const double *kY, int n, const double *patterns, const double *kt,
                               double f, int kind, const double *prop_mats, double diskR, int includes_axion,
                               const double *emit, double *fields_out) {
    coords_t c = {nx, ny, X, Y, kX, kY};
    modes_t md = {0, 0, n, (const cx *)patterns, (const cx *)kt};
    size_t K = (size_t)nx * ny;
    int N2 = 2 * n;
    double lambda = bfo_wavelength(f);
    cx *sol = (cx *)malloc(sizeof(cx) * (size_t)N2), *nsol = (cx *)malloc(sizeof(cx) * (size_t)N2);
    cx *P = (cx *)malloc(sizeof(cx) * (size_t)n * n), *Bm = (cx *)malloc(sizeof(cx) * (size_t)N2 * N2);
    cx *out = (cx *)fields_out;
    /* fields_regions[r,:,:] = solution: the 2n vector fills the 2 x n slice in
     * column-major order, element (d, m) = solution[d + 2m]; out is [nreg][2][n] */
#define PUT(r, v) do { for (int d_ = 0; d_ < 2; d_++) for (int m_ = 0; m_ < n; m_++) \
        out[((size_t)(r) * 2 + d_) * n + m_] = (v)[d_ + 2 * m_]; } while (0)
    for (int i = 0; i < n; i++) { sol[i] = cload(input, i); sol[n + i] = cload(reflected, i); }
    PUT(nreg - 1, sol);
    fft2 pl;
    fft2_init(&pl, nx, ny);
    for (int s = 1; s <= nreg - 1; s++) {
        int ir = nreg - s, ir1 = nreg - s - 1;
        cx e_ir = cload(eps, ir), e_ir1 = cload(eps, ir1);
        if (prop_mats)
            memcpy(P, (const cx *)prop_mats + (size_t)ir * n * n, sizeof(cx) * (size_t)n * n);
        else
            propagation_matrix_c(distance[ir], diskR, e_ir, tilt_x[ir], tilt_y[ir],
                                 surfaces ? surfaces + 2 * (size_t)ir * K : NULL, lambda, &c, &pl, &md, kind, 0, P);
        cx nl = csqrt_j(e_ir), nr = csqrt_j(e_ir1);
        boundary_matrix(nl, nr, P, n, Bm);
        for (int i = 0; i < N2; i++) {
            cx acc = C(0.0, 0.0);
            for (int k = 0; k < N2; k++) acc = cadd(acc, cmul(Bm[i + (size_t)N2 * k], sol[k]));
            nsol[i] = acc;
        }
        memcpy(sol, nsol, sizeof(cx) * (size_t)N2);
        if (includes_axion) {
            cx S = rmul(0.5, csub(cdiv(C(1.0, 0.0), cmul(nr, nr)), cdiv(C(1.0, 0.0), cmul(nl, nl))));
            for (int i = 0; i < N2; i++) sol[i] = cadd(sol[i], cmul(S, cload(emit, i % n)));
        }
        PUT(ir1, sol);
    }
#undef PUT
    fft2_free(&pl);
    free(sol); free(nsol); free(P); free(Bm);
    return 0;
}
