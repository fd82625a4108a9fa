// bf_beam.hip -- antenna-side helpers: Gaussian beam profile and field overlap.
//
//   gauss_profile       src/beam_gaussian.jl:24-37
//   unnormalized_match  src/beam_coupling.jl:13-15
#include "bf_internal.h"

#include <cmath>

namespace bf {

struct GaussParams {
    double k, Rz, psi, z, amp0, om2;
};

// omega0/omega(z) .* exp.(-radius.^2/omega(z)^2) .* exp.(-1im .* (k*z .+ k*radius.^2 ./(2 R(z)) .- Psi(z)))
__global__ void k_gauss_profile(zd* out, int nx, int ny, const double* __restrict__ X, const double* __restrict__ Y,
                                GaussParams p) {
#pragma clang fp contract(off)
    const size_t K = (size_t)nx * ny;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < K; i += (size_t)gridDim.x * blockDim.x) {
        const double x = X[i % nx], y = Y[i / nx];
        const double r = sqrt(x * x + y * y);
        const double r2 = r * r;
        const double ph = p.k * p.z + p.k * r2 / (2.0 * p.Rz) - p.psi;
        const zd e = cexp_j(cmk<double>(0.0 * ph, -ph));  // -1im * ph
        const double a = p.amp0 * exp(-r2 / p.om2);
        out[i] = cmk<double>(a * e.re, a * e.im);
    }
}

void gauss_profile_dev(bf_ctx* ctx, const bf_coords* co, double z, double omega0, double f, double* out) {
    hipStream_t s = ctx->stream;
    Grid g;
    g.upload(co, s);
    const size_t K = g.K();
    const double lambda = 299792458.0 / f;
    const double k = 2.0 * M_PI / lambda;
    const double zR = M_PI * omega0 * omega0 / lambda;
    const double om = omega0 * std::sqrt(1.0 + (z / zR) * (z / zR));
    GaussParams p;
    p.k = k;
    p.Rz = z * (1.0 + (zR / z) * (zR / z));
    p.psi = std::atan(z / zR);
    p.z = z;
    p.amp0 = omega0 / om;
    p.om2 = om * om;
    DevBuf d(sizeof(zd) * K);
    hipLaunchKernelGGL(k_gauss_profile, dim3((unsigned)std::min<size_t>((K + 255) / 256, 4096)), dim3(256), 0, s,
                       d.as<zd>(), g.nx, g.ny, g.X.as<double>(), g.Y.as<double>(), p);
    BF_CHECK_LAUNCH();
    download(out, d.p, sizeof(zd) * K, s);
    BF_HIP(hipStreamSynchronize(s));
}

// sum(Eout .* conj(Eantenna)) for `nb` fields against one antenna pattern:
// per-block partial sums in a fixed order, then a fixed-order sum of the
// partials (deterministic; differs from Julia's pairwise order by rounding).
constexpr int MATCH_BLOCKS = 256;

__global__ void k_match_partial(const zd* __restrict__ E, const zd* __restrict__ A, size_t n, zd* partial) {
    __shared__ zd red[256];
    const int b = blockIdx.y;
    const zd* e = E + (size_t)b * n;
    zd acc = cmk<double>(0.0, 0.0);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        acc = cadd(acc, cmulc(e[i], A[i]));
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = cadd(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[(size_t)b * gridDim.x + blockIdx.x] = red[0];
}

__global__ void k_match_final(const zd* __restrict__ partial, int nparts, double dx, double dy, zd* out) {
    const int b = blockIdx.x;
    if (threadIdx.x != 0) return;
    zd s = cmk<double>(0.0, 0.0);
    for (int q = 0; q < nparts; q++) s = cadd(s, partial[(size_t)b * nparts + q]);
    s = cmk<double>(s.re * dx, s.im * dx);
    out[b] = cmk<double>(s.re * dy, s.im * dy);
}

void unnormalized_match_dev(bf_ctx* ctx, int nb, int64_t n, const double* Eout, const double* Eant, double dx,
                            double dy, double* out) {
    hipStream_t s = ctx->stream;
    DevBuf dE(sizeof(zd) * (size_t)n * nb), dA(sizeof(zd) * (size_t)n), dp(sizeof(zd) * (size_t)MATCH_BLOCKS * nb),
        dout(sizeof(zd) * nb);
    upload(dE.p, Eout, sizeof(zd) * (size_t)n * nb, s);
    upload(dA.p, Eant, sizeof(zd) * (size_t)n, s);
    hipLaunchKernelGGL(k_match_partial, dim3(MATCH_BLOCKS, nb), dim3(256), 0, s, dE.as<zd>(), dA.as<zd>(), (size_t)n,
                       dp.as<zd>());
    BF_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_match_final, dim3(nb), dim3(64), 0, s, dp.as<zd>(), MATCH_BLOCKS, dx, dy, dout.as<zd>());
    BF_CHECK_LAUNCH();
    download(out, dout.p, sizeof(zd) * nb, s);
    BF_HIP(hipStreamSynchronize(s));
}

}  // namespace bf
