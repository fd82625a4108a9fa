// bf_analytic1d.hip -- the 1D analytical disk model (src/Analytical1D.py) on gfx950.
//
//   transform_surface   src/Analytical1D.py:5-67
//   disk_system         src/Analytical1D.py:82-131
//
// The model walks the surfaces of a booster from the mirror outwards with two
// complex numbers per frequency point: the reflectivity gamma in front of the
// surface (ray-tracing form, or the impedance transformation) and the
// outgoing axion amplitude.  One thread runs the whole recurrence of one
// (candidate geometry, frequency) point; a launch covers every point of a
// spacing sweep.  FP64 throughout.
#include "bf_internal.h"

#include <algorithm>
#include <cmath>

namespace bf {

// complex quotient by Smith's method (R. L. Smith, CACM 5(8), 1962)
__host__ __device__ inline zd cdiv_smith(zd a, zd b) {
    if (fabs(b.re) >= fabs(b.im)) {
        const double r = b.im / b.re, d = b.re + b.im * r;
        return cmk<double>((a.re + a.im * r) / d, (a.im - a.re * r) / d);
    }
    const double r = b.re / b.im, d = b.re * r + b.im;
    return cmk<double>((a.re * r + a.im) / d, (a.im * r - a.re) / d);
}

__host__ __device__ inline zd cfrom(double x) { return cmk<double>(x, 0.0); }

// numpy's complex quotient (Smith's method with a reciprocal), operation for
// operation: the propagation phases below are formed the way the reference's
// numpy expressions form them.  Inside the model's metal mirror (eps = 1e20,
// 0.1 m) the phase is ~1e12 rad, where one rounding step is ~1e-4 rad.
__host__ __device__ inline zd cdiv_np(zd a, zd b) {
#pragma clang fp contract(off)
    if (fabs(b.re) >= fabs(b.im)) {
        const double rat = b.im / b.re, scl = 1.0 / (b.re + b.im * rat);
        return cmk<double>((a.re + a.im * rat) * scl, (a.im - a.re * rat) * scl);
    }
    const double rat = b.re / b.im, scl = 1.0 / (b.im + b.re * rat);
    return cmk<double>((a.re * rat + a.im) * scl, (a.im * rat - a.re) * scl);
}

// exp(z) for a phase that may be huge (~1e12 rad inside the model's metal):
// the angle is first reduced exactly by 2 pi, split as C1 + C2 (C1 = fl(2 pi),
// C2 = 2 pi - C1): k C1 is exact inside an FMA and the remainder fits in a
// double, so the angle passed to sincos is correct to ~1e-15 for |angle| up to
// 2^52 (the device's own reduction is good to ~1e-16 of the angle, 1e-4 rad here).
__host__ __device__ inline zd cexp_phase(zd z) {
    double t = z.im;
    if (fabs(t) > 1e5) {
        const double C1 = 6.283185307179586, C2 = 2.4492935982947064e-16;
        const double k = rint(t * (1.0 / C1));
        t = fma(-k, C1, t);
        t = fma(-k, C2, t);
    }
    const double er = exp(z.re);
    double s, c;
#ifdef __HIP_DEVICE_COMPILE__
    sincos(t, &s, &c);
#else
    s = sin(t);
    c = cos(t);
#endif
    return cmk<double>(er * c, er * s);
}

// exp(-1j*k*pi*l/wavel) as numpy evaluates it, k = 4 (reflectivity) or 2 (axion):
// wavel = c / (freq * sqrt(eps_m)) in real arithmetic when eps_m is a real
// number in the reference call (m_real), in complex arithmetic otherwise.
__host__ __device__ inline zd phase_np(double k, double freq, zd eps_m, bool m_real, double l) {
#pragma clang fp contract(off)
    const zd num = cmk<double>(-0.0 * l, (-k * M_PI) * l);  // ((-1j*k)*pi)*l
    zd q;
    if (m_real) {
        const double wl = 299792458.0 / (freq * sqrt(eps_m.re));
        q = cdiv_np(num, cmk<double>(wl, 0.0));
    } else {
        const zd sm = csqrt_j(eps_m);
        const zd fs = cmk<double>(freq * sm.re - 0.0 * sm.im, freq * sm.im + 0.0 * sm.re);
        const zd wl = cdiv_np(cmk<double>(299792458.0, 0.0), fs);
        q = cdiv_np(num, wl);
    }
    return cexp_phase(q);
}

// One surface (src/Analytical1D.py:5-67): the medium behind it has eps_i, the
// one in front eps_m and thickness l.  `ax` (when non-null) is the outgoing
// axion amplitude, updated in place.
__host__ __device__ inline zd transform_surface_1d(double freq, zd gamma, zd* ax, zd eps_i, zd eps_m, bool m_real,
                                                   double l, double surfaceloss, bool ray) {
    const zd si = csqrt_j(eps_i), sm = csqrt_j(eps_m);
    const zd one = cfrom(1.0);
    zd r_s = cmk<double>(0.0, 0.0);
    const zd r_int = gamma;
    const double g = 1.0 - surfaceloss;
    if (ray) {
        r_s = cdiv_smith(csub(sm, si), cadd(sm, si));
        // gamma = g r_s + g^2 (1 + r_s)(1 - r_s) r_int / (1 + r_int r_s g)
        const zd num = cmul(rmul(g * g, cmul(cadd(one, r_s), csub(one, r_s))), r_int);
        const zd den = cadd(one, rmul(g, cmul(r_int, r_s)));
        gamma = cadd(rmul(g, r_s), cdiv_smith(num, den));
    } else {
        // impedance transformation: z = (1 + gamma)/(1 - gamma) * Z_i/Z_m, Z = 1/sqrt(eps)
        zd z = cdiv_smith(cadd(one, gamma), csub(one, gamma));
        z = cmul(z, cdiv_smith(cdiv_smith(one, si), cdiv_smith(one, sm)));
        gamma = rmul(1.0 - surfaceloss, cdiv_smith(csub(z, one), cadd(z, one)));
    }
    // bidirectional propagation exp(-i 4 pi l / wavel), wavel = c / (f sqrt(eps_m))
    gamma = cmul(gamma, phase_np(4.0, freq, eps_m, m_real, l));
    if (ax) {
        const zd den = cadd(cmul(eps_i, sm), cmul(eps_m, si));
        zd ax_i = cdiv_smith(cmul(si, csub(one, cdiv_smith(eps_m, eps_i))), den);
        const zd ax_m = cdiv_smith(cmul(sm, csub(one, cdiv_smith(eps_i, eps_m))), den);
        // the new axion field inside, through the surface
        ax_i = cmul(ax_i, cdiv_smith(rmul(g, cmul(r_int, csub(one, r_s))), cadd(one, rmul(g, cmul(r_int, r_s)))));
        // the present outgoing field through the surface
        zd a = cmul(*ax, cdiv_smith(rmul(g, csub(one, r_s)), cadd(one, rmul(g, cmul(r_int, r_s)))));
        a = cadd(a, cadd(ax_i, ax_m));
        *ax = cmul(a, phase_np(2.0, freq, eps_m, m_real, l));
    }
    return gamma;
}

struct Disk1D {
    zd eps_disk, eps_gap;
    double thickness;
    int num_disk, mirror, has_loss;
};

// disk_system (src/Analytical1D.py:82-131) for every (candidate, frequency) point:
// spacings [ncand][num_disk + 1], surfaceloss [2 num_disk + 1] (has_loss)
__global__ void k_disk_system_1d(const double* __restrict__ freqs, int nfreq, int ncand, Disk1D d,
                                 const double* __restrict__ spacings, const double* __restrict__ sloss, zd* gamma_out,
                                 zd* ax_out) {
    const size_t total = (size_t)ncand * nfreq;
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < total; p += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(p / nfreq);
        const double f = freqs[p % nfreq];
        const double* sp = spacings + (size_t)c * (d.num_disk + 1);
        zd ax = cmk<double>(0.0, 0.0);
        zd gm = cmk<double>(0.0, 0.0);
        const zd e_air = cfrom(1.0), e_metal = cfrom(1e20);
        // the metal's 1e20 is a real number in the reference call; the gap and disk
        // permittivities come from lossy_eps as complex numbers
        if (d.mirror) gm = transform_surface_1d(f, gm, &ax, e_air, e_metal, true, 100e-3, 0.0, true);
        gm = transform_surface_1d(f, gm, &ax, d.mirror ? e_metal : e_air, d.eps_gap, false, sp[0],
                                  d.has_loss ? sloss[0] : 0.0, true);
        for (int i = 0; i < d.num_disk; i++) {
            gm = transform_surface_1d(f, gm, &ax, d.eps_gap, d.eps_disk, false, d.thickness,
                                      d.has_loss ? sloss[2 * i + 1] : 0.0, true);
            gm = transform_surface_1d(f, gm, &ax, d.eps_disk, d.eps_gap, false, sp[i + 1],
                                      d.has_loss ? sloss[2 * i + 2] : 0.0, true);
        }
        gamma_out[p] = gm;
        ax_out[p] = ax;
    }
}

// transform_surface over arrays of frequencies (ax in and out when given)
__global__ void k_transform_surface_1d(const double* __restrict__ freqs, int n, const zd* __restrict__ gin, zd* ax,
                                       zd eps_i, zd eps_m, int m_real, double l, double surfaceloss, int ray,
                                       zd* gout) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x)
        gout[p] = transform_surface_1d(freqs[p], gin[p], ax ? ax + p : nullptr, eps_i, eps_m, m_real != 0, l,
                                       surfaceloss, ray != 0);
}

void disk_system_1d_dev(bf_ctx* ctx, const double* freqs, int nfreq, int ncand, int num_disk, const double* spacings,
                        double thickness, const double* eps_disk, const double* eps_gap, int mirror,
                        const double* surfaceloss, double* gamma, double* axion) {
    hipStream_t s = ctx->stream;
    const size_t total = (size_t)ncand * nfreq;
    DevBuf df(sizeof(double) * nfreq), dsp(sizeof(double) * (size_t)ncand * (num_disk + 1)),
        dsl(sizeof(double) * (2 * (size_t)num_disk + 1)), dg(sizeof(zd) * total), da(sizeof(zd) * total);
    upload(df.p, freqs, sizeof(double) * nfreq, s);
    upload(dsp.p, spacings, sizeof(double) * (size_t)ncand * (num_disk + 1), s);
    if (surfaceloss) upload(dsl.p, surfaceloss, sizeof(double) * (2 * (size_t)num_disk + 1), s);
    Disk1D d;
    d.eps_disk = cmk<double>(eps_disk[0], eps_disk[1]);
    d.eps_gap = cmk<double>(eps_gap[0], eps_gap[1]);
    d.thickness = thickness;
    d.num_disk = num_disk;
    d.mirror = mirror;
    d.has_loss = surfaceloss != nullptr;
    const unsigned blocks = (unsigned)std::max<size_t>(1, std::min<size_t>((total + 255) / 256, 65535));
    {
        TimedLaunch tl(ctx, "disk_system_1d", 0.0);
        hipLaunchKernelGGL(k_disk_system_1d, dim3(blocks), dim3(256), 0, s, df.as<double>(), nfreq, ncand, d,
                           dsp.as<double>(), dsl.as<double>(), dg.as<zd>(), da.as<zd>());
        BF_CHECK_LAUNCH();
    }
    download(gamma, dg.p, sizeof(zd) * total, s);
    download(axion, da.p, sizeof(zd) * total, s);
    BF_HIP(hipStreamSynchronize(s));
}

void transform_surface_1d_dev(bf_ctx* ctx, const double* freqs, int n, const double* gamma_in, double* axion,
                              const double* eps_i, const double* eps_m, int m_real, double l, double surfaceloss,
                              int ray, double* gamma_out) {
    hipStream_t s = ctx->stream;
    DevBuf df(sizeof(double) * n), dgi(sizeof(zd) * n), dgo(sizeof(zd) * n), dax;
    upload(df.p, freqs, sizeof(double) * n, s);
    upload(dgi.p, gamma_in, sizeof(zd) * n, s);
    if (axion) {
        dax.alloc(sizeof(zd) * n);
        upload(dax.p, axion, sizeof(zd) * n, s);
    }
    const unsigned blocks = (unsigned)std::max(1, std::min((n + 255) / 256, 65535));
    hipLaunchKernelGGL(k_transform_surface_1d, dim3(blocks), dim3(256), 0, s, df.as<double>(), n, dgi.as<zd>(),
                       dax.as<zd>(), cmk<double>(eps_i[0], eps_i[1]), cmk<double>(eps_m[0], eps_m[1]), m_real, l,
                       surfaceloss, ray, dgo.as<zd>());
    BF_CHECK_LAUNCH();
    download(gamma_out, dgo.p, sizeof(zd) * n, s);
    if (axion) download(axion, dax.p, sizeof(zd) * n, s);
    BF_HIP(hipStreamSynchronize(s));
}

}  // namespace bf
