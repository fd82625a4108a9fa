// bf_sweep.hip -- the disk-position sweep around the transformer
// (examples/transformer_optim_utilities.jl:11-160) on gfx950.
//
//   propagation_matrices_grid   calc_propagation_matrices_grid :11-52
//   SpacingItp                  construct_prop_matrix_interpolation :58-74,
//                               interpolate_prop_matrix :79-91,
//                               calc_boostfactor_modes / calc_modes :99-125
//
// The interpolation is Interpolations.jl's BSpline(Cubic(Natural(OnCell())))
// along the spacing axis (Natural = Line: zero second derivative at the outer
// cell edges x = 1/2 and x = n + 1/2 of the 1-based grid index x), scaled to
// the uniform spacing grid.  With padded coefficients c_0 .. c_{n+1}, the spline
//     s(x) = sum_q w_q(d) c_{i-1+q},  i = clamp(floor(x), 1, n-1),  d = x - i,
//     w_0 = (1-d)^3/6, w_1 = 2/3 - d^2 + d^3/2, w_2 = 2/3 - (1-d)^2 + (1-d)^3/2, w_3 = d^3/6
// interpolates y_1..y_n (c_{i-1} + 4 c_i + c_{i+1} = 6 y_i); the two edge
// conditions s''(1/2) = s''(n+1/2) = 0 evaluated on the first and last cells'
// polynomials read 3c_0 - 7c_1 + 5c_2 - c_3 = 0 and -c_{n-2} + 5c_{n-1} - 7c_n +
// 3c_{n+1} = 0.  Eliminating c_0 and c_{n+1} with the first and last
// interpolation rows leaves the diagonally dominant tridiagonal system
//     -3 c_1 + c_2 = y_2 - 3 y_1,
//     c_{i-1} + 4 c_i + c_{i+1} = 6 y_i              (1 < i < n),
//     c_{n-1} - 3 c_n = y_{n-1} - 3 y_n,
// solved per matrix element by the Thomas algorithm (one thread per element,
// consecutive threads on consecutive elements), then
// c_0 = 6 y_1 - 4 c_1 - c_2 and c_{n+1} = 6 y_n - 4 c_n - c_{n-1}.
#include "bf_modematch.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace bf {

// ---------------------------------------------------------------------------
// propagation-matrix grid (:11-52)
// ---------------------------------------------------------------------------
void propagation_matrices_grid_dev(bf_ctx* ctx, const bf_setup* st, const bf_coords* co, const bf_modes* modes,
                                   const double* spacing, int ns, const double* tgx, int ntx, const double* tgy,
                                   int nty, const double* freqs, int nf, int kind, double diskR, double* out) {
    hipStream_t s = ctx->stream;
    Grid g;
    g.upload(co, s);
    const size_t K = g.K();
    ModesDev md;
    md.upload(modes, K, s);
    ModeCache cache;
    const int n = md.n, nreg = st->nreg, nsub = nreg / 2;
    const size_t nn = (size_t)n * n;
    // gaps: 0-based regions 1, 3, ..., nreg-1 (ind_gap = 2:2:n_region); solids: 0, 2, ..., nreg-2
    std::vector<zd> eps_all(reinterpret_cast<const zd*>(st->eps), reinterpret_cast<const zd*>(st->eps) + nreg);
    std::vector<double> dist[2], tx[2], ty[2];
    std::vector<zd> eps[2];
    for (int p = 0; p < 2; p++)
        for (int q = 0; q < nsub; q++) {
            const int r = 2 * q + (p == 0 ? 1 : 0);
            dist[p].push_back(st->distance[r]);
            tx[p].push_back(st->tilt_x[r]);
            ty[p].push_back(st->tilt_y[r]);
            zd e = eps_all[r];
            if (e.im == 0.0) e.im = -0.0;  // the sub-setups pass through SeedSetupBoundaries (src/boost_fractor.jl:96)
            eps[p].push_back(e);
        }
    DevBuf surf[2];
    std::vector<bool> flat[2];
    for (int p = 0; p < 2; p++) {
        flat[p].assign(nsub, true);
        if (!st->surfaces) continue;
        const zd* sf = reinterpret_cast<const zd*>(st->surfaces);
        std::vector<zd> sub((size_t)nsub * K);
        for (int q = 0; q < nsub; q++) {
            const int r = 2 * q + (p == 0 ? 1 : 0);
            std::memcpy(&sub[(size_t)q * K], sf + (size_t)r * K, sizeof(zd) * K);
            for (size_t i = 0; i < K; i++)
                if (sub[(size_t)q * K + i].re != 0.0 || sub[(size_t)q * K + i].im != 0.0) {
                    flat[p][q] = false;
                    break;
                }
        }
        surf[p].alloc(sizeof(zd) * K * nsub);
        upload(surf[p].p, sub.data(), sizeof(zd) * K * nsub, s);
    }
    std::vector<int> regs(nsub);
    for (int q = 0; q < nsub; q++) regs[q] = q;
    const size_t per = (size_t)nf * nsub * nn;  // matrices of one sub-setup at all frequencies
    DevBuf dP(sizeof(zd) * per);
    std::vector<zd> solid(per), gap(per);
    // solids: once (calc_propagation_matrices(sbdry_solid) per frequency)
    propagation_matrices_core(ctx, g, md, kind, diskR, nsub, dist[1].data(), eps[1].data(), tx[1].data(),
                              ty[1].data(), surf[1].as<zd>(), flat[1], freqs, nf, regs, dP.as<zd>(), &cache);
    download(solid.data(), dP.p, sizeof(zd) * per, s);
    BF_HIP(hipStreamSynchronize(s));
    zd* o = reinterpret_cast<zd*>(out);
    std::vector<double> dc(nsub), txc(nsub), tyc(nsub);
    for (int is = 0; is < ns; is++)
        for (int ix = 0; ix < ntx; ix++)
            for (int iy = 0; iy < nty; iy++) {
                for (int q = 0; q < nsub; q++) {
                    dc[q] = dist[0][q] + spacing[is];
                    txc[q] = tx[0][q] + tgx[ix];
                    tyc[q] = ty[0][q] + tgy[iy];
                }
                propagation_matrices_core(ctx, g, md, kind, diskR, nsub, dc.data(), eps[0].data(), txc.data(),
                                          tyc.data(), surf[0].as<zd>(), flat[0], freqs, nf, regs, dP.as<zd>(),
                                          &cache);
                download(gap.data(), dP.p, sizeof(zd) * per, s);
                BF_HIP(hipStreamSynchronize(s));
                zd* blk = o + (((size_t)is * ntx + ix) * nty + iy) * nf * nreg * nn;
                for (int f = 0; f < nf; f++)
                    for (int q = 0; q < nsub; q++) {
                        const size_t src = ((size_t)f * nsub + q) * nn;
                        std::memcpy(blk + ((size_t)f * nreg + 2 * q + 1) * nn, &gap[src], sizeof(zd) * nn);
                        std::memcpy(blk + ((size_t)f * nreg + 2 * q) * nn, &solid[src], sizeof(zd) * nn);
                    }
            }
}

// ---------------------------------------------------------------------------
// cubic B-spline over the spacing axis
// ---------------------------------------------------------------------------
// coef: [ns + 2][S] (padded index slowest); on entry rows 1..ns hold the data.
// cp[i], rinv[i] (i = 1..ns): the Thomas algorithm's modified upper diagonal and
// reciprocal pivots of the tridiagonal system above (the same for every series).
__global__ void k_spline_prefilter(zd* coef, size_t S, int ns, const double* __restrict__ cp,
                                   const double* __restrict__ rinv) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < S; e += (size_t)gridDim.x * blockDim.x) {
        zd* c = coef + e;
        const zd y1 = c[S], y2 = c[2 * S];
        const zd yn = c[(size_t)ns * S], yn1 = c[(size_t)(ns - 1) * S];
        // forward sweep: rows 1..ns; the right-hand side of row i needs y_i (rows 1 and ns also a neighbour)
        zd dp = cmk<double>((y2.re - 3.0 * y1.re) * rinv[1], (y2.im - 3.0 * y1.im) * rinv[1]);
        c[S] = dp;
        for (int i = 2; i < ns; i++) {
            const zd y = c[(size_t)i * S];
            const double r = rinv[i];
            dp = cmk<double>((6.0 * y.re - dp.re) * r, (6.0 * y.im - dp.im) * r);
            c[(size_t)i * S] = dp;
        }
        {
            const double r = rinv[ns];  // row ns (ns >= 2)
            dp = cmk<double>((yn1.re - 3.0 * yn.re - dp.re) * r, (yn1.im - 3.0 * yn.im - dp.im) * r);
            c[(size_t)ns * S] = dp;
        }
        // back substitution: x_ns = dp_ns, x_i = dp_i - cp_i x_{i+1}
        zd x = dp;
        const zd xn = dp;
        zd xn1 = dp, x2 = dp;  // x_{ns-1} and x_2 (both x_ns's neighbours when ns == 2)
        for (int i = ns - 1; i >= 1; i--) {
            const zd d = c[(size_t)i * S];
            x = cmk<double>(d.re - cp[i] * x.re, d.im - cp[i] * x.im);
            c[(size_t)i * S] = x;
            if (i == ns - 1) xn1 = x;
            if (i == 2) x2 = x;
        }
        // x = c_1
        c[0] = cmk<double>(6.0 * y1.re - 4.0 * x.re - x2.re, 6.0 * y1.im - 4.0 * x.im - x2.im);
        c[(size_t)(ns + 1) * S] =
            cmk<double>(6.0 * yn.re - 4.0 * xn.re - xn1.re, 6.0 * yn.im - 4.0 * xn.im - xn1.im);
    }
}

struct SplinePoint {
    double w[4];
    int p0;  // padded index of w[0]'s coefficient
    int pad;
};

// P[c][f][r] = sum_q w_q coef[p0 + q][f][r] for every candidate c (one thread per element)
__global__ void k_spline_eval(const zd* __restrict__ coef, size_t S, int nfreg, int nreg, int nn,
                              const SplinePoint* __restrict__ pts, size_t total, zd* P) {
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t m = t / nn, e = t % nn;
        const size_t c = m / nfreg, fr = m % nfreg;  // fr = f * nreg + r
        const int r = (int)(fr % nreg);
        const SplinePoint pt = pts[c * nreg + r];
        const zd* src = coef + (size_t)pt.p0 * S + fr * nn + e;
        double re = 0.0, im = 0.0;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const zd v = src[(size_t)q * S];
            re += pt.w[q] * v.re;
            im += pt.w[q] * v.im;
        }
        P[t] = cmk<double>(re, im);
    }
}

bf_spacing_itp* spacing_itp_create(bf_ctx* ctx, const double* grid, int nreg, int nf, int ns, int n, double first,
                                   double step) {
    hipStream_t s = ctx->stream;
    auto* it = new bf_spacing_itp;
    try {
        it->ctx = ctx;
        it->nreg = nreg;
        it->nf = nf;
        it->ns = ns;
        it->n = n;
        it->first = first;
        it->step = step;
        const size_t S = (size_t)nf * nreg * n * n;
        it->coef.alloc(sizeof(zd) * S * (ns + 2));
        upload(it->coef.as<zd>() + S, grid, sizeof(zd) * S * ns, s);
        // Thomas factors of the tridiagonal system (1-based rows 1..ns)
        std::vector<double> a(ns + 1, 1.0), b(ns + 1, 4.0), c(ns + 1, 1.0), cp(ns + 1, 0.0), rinv(ns + 1, 0.0);
        b[1] = -3.0;
        c[1] = 1.0;
        a[1] = 0.0;
        b[ns] = -3.0;
        a[ns] = 1.0;
        c[ns] = 0.0;
        double prev = 0.0;
        for (int i = 1; i <= ns; i++) {
            const double den = b[i] - a[i] * prev;
            rinv[i] = 1.0 / den;
            cp[i] = c[i] / den;
            prev = cp[i];
        }
        DevBuf dcp(sizeof(double) * (ns + 1)), dri(sizeof(double) * (ns + 1));
        upload(dcp.p, cp.data(), sizeof(double) * (ns + 1), s);
        upload(dri.p, rinv.data(), sizeof(double) * (ns + 1), s);
        const unsigned blocks = (unsigned)std::max<size_t>(1, std::min<size_t>((S + 255) / 256, 8192));
        hipLaunchKernelGGL(k_spline_prefilter, dim3(blocks), dim3(256), 0, s, it->coef.as<zd>(), S, ns,
                           dcp.as<double>(), dri.as<double>());
        BF_CHECK_LAUNCH();
        BF_HIP(hipStreamSynchronize(s));
    } catch (...) {
        delete it;
        throw;
    }
    return it;
}

void spacing_itp_destroy(bf_spacing_itp* it) { delete it; }

void spacing_itp_bounds(const bf_spacing_itp* it, double* lo, double* hi) {
    *lo = it->first - 0.5 * it->step;
    *hi = it->first + (it->ns - 1) * it->step + 0.5 * it->step;
}

// spline weights of one spacing shift (x in grid units, 1-based)
static SplinePoint spline_point(const bf_spacing_itp* it, double shift) {
    double lo, hi;
    spacing_itp_bounds(it, &lo, &hi);
    if (!(shift >= lo && shift <= hi))  // also rejects NaN
        throw Error(BF_ERR_ARG, "BoundsError: spacing shift " + std::to_string(shift) +
                                    " outside the interpolation bounds [" + std::to_string(lo) + ", " +
                                    std::to_string(hi) + "]");
    const double x = (shift - it->first) / it->step + 1.0;
    double i = std::floor(x);
    i = std::min(std::max(i, 1.0), (double)(it->ns - 1));
    const double d = x - i, e = 1.0 - d;
    SplinePoint p;
    p.w[0] = e * e * e / 6.0;
    p.w[1] = 2.0 / 3.0 - d * d + d * d * d / 2.0;
    p.w[2] = 2.0 / 3.0 - e * e + e * e * e / 2.0;
    p.w[3] = d * d * d / 6.0;
    p.p0 = (int)i - 1;
    p.pad = 0;
    return p;
}

// interpolate_prop_matrix (:79-91): region r (1-based) is evaluated at 0 for odd r
// and for r = Nreg, at dist_shift[r/2] otherwise.  ndisk = (Nreg - 2)/2 shifts per candidate.
static void spline_points(const bf_spacing_itp* it, int c0, int nc, const double* shifts,
                          std::vector<SplinePoint>& pts) {
    const int ndisk = (it->nreg - 2) / 2;
    pts.clear();
    for (int c = c0; c < c0 + nc; c++)
        for (int r = 1; r <= it->nreg; r++) {
            double x = 0.0;
            if (r % 2 == 0 && r != it->nreg) x = shifts[(size_t)c * ndisk + r / 2 - 1];
            pts.push_back(spline_point(it, x));
        }
}

static void spline_eval(bf_spacing_itp* it, const std::vector<SplinePoint>& pts, int nc, zd* P) {
    hipStream_t s = it->ctx->stream;
    const size_t nn = (size_t)it->n * it->n, S = (size_t)it->nf * it->nreg * nn;
    DevBuf dp(sizeof(SplinePoint) * pts.size());
    upload(dp.p, pts.data(), sizeof(SplinePoint) * pts.size(), s);
    const size_t total = (size_t)nc * S;
    const unsigned blocks = (unsigned)std::max<size_t>(1, std::min<size_t>((total + 255) / 256, 16384));
    {
        TimedLaunch tl(it->ctx, "spline_eval", 8.0 * (double)total);
        hipLaunchKernelGGL(k_spline_eval, dim3(blocks), dim3(256), 0, s, it->coef.as<zd>(), S, it->nf * it->nreg,
                           it->nreg, (int)nn, dp.as<SplinePoint>(), total, P);
        BF_CHECK_LAUNCH();
    }
    BF_HIP(hipStreamSynchronize(s));
}

// candidates per chunk so that P and inv(P) stay within `budget` bytes
static int cand_chunk(const bf_spacing_itp* it, int ncand) {
    const size_t budget = opts().batch_bytes > 0 ? (size_t)opts().batch_bytes : BF_DEFAULT_BATCH_BYTES;
    const size_t per = (size_t)it->nf * it->nreg * it->n * it->n * sizeof(zd) * 2;
    return (int)std::max<size_t>(1, std::min<size_t>((size_t)ncand, budget / per));
}

void spacing_itp_matrices(bf_spacing_itp* it, int ncand, const double* shifts, double* out) {
    hipStream_t s = it->ctx->stream;
    const size_t per = (size_t)it->nf * it->nreg * it->n * it->n;
    const int cb = cand_chunk(it, ncand);
    DevBuf P(sizeof(zd) * per * cb);
    std::vector<SplinePoint> pts;
    for (int c0 = 0; c0 < ncand; c0 += cb) {
        const int nc = std::min(cb, ncand - c0);
        spline_points(it, c0, nc, shifts, pts);
        spline_eval(it, pts, nc, P.as<zd>());
        download(reinterpret_cast<zd*>(out) + (size_t)c0 * per, P.p, sizeof(zd) * per * nc, s);
        BF_HIP(hipStreamSynchronize(s));
    }
}

// calc_boostfactor_modes / calc_modes (:99-125) with interpolate_prop_matrix for
// every candidate: transformer(...; propagation_matrices=...) per (candidate, frequency).
void spacing_itp_transformer(bf_spacing_itp* it, int ncand, const double* shifts, const double* eps_in,
                             const double* emit, const double* reflect, double* boost, double* refl) {
    bf_ctx* ctx = it->ctx;
    hipStream_t s = ctx->stream;
    const int n = it->n, nreg = it->nreg, nf = it->nf;
    const size_t nn = (size_t)n * n, per = (size_t)nf * nreg * nn;
    std::vector<zd> eps(reinterpret_cast<const zd*>(eps_in), reinterpret_cast<const zd*>(eps_in) + nreg);
    for (auto& z : eps)
        if (std::isnan(z.re) || std::isnan(z.im)) z = cmk<double>(1e30, 0.0);  // bdry.eps[isnan] .= 1e30
    std::vector<ChainStep> steps;
    append_chain_steps(eps, nreg, steps);  // the same region sequence for every job
    const int cb = cand_chunk(it, ncand);
    DevBuf P(sizeof(zd) * per * cb), Pinv(sizeof(zd) * per * cb), de0(sizeof(zd) * n), dref, dboost(sizeof(zd) * n * nf * cb),
        drefl;
    upload(de0.p, emit, sizeof(zd) * n, s);
    if (reflect) {
        dref.alloc(sizeof(zd) * n);
        drefl.alloc(sizeof(zd) * n * nf * cb);
        upload(dref.p, reflect, sizeof(zd) * n, s);
    }
    std::vector<SplinePoint> pts;
    for (int c0 = 0; c0 < ncand; c0 += cb) {
        const int nc = std::min(cb, ncand - c0);
        spline_points(it, c0, nc, shifts, pts);
        spline_eval(it, pts, nc, P.as<zd>());
        std::vector<int> sel;
        std::vector<ChainJob> jobs;
        for (int c = 0; c < nc; c++)
            for (int f = 0; f < nf; f++) {
                const int64_t base = ((int64_t)c * nf + f) * nreg;
                for (int r = 1; r < nreg; r++) sel.push_back((int)(base + r));
                ChainJob j;
                j.step0 = 0;
                j.nsteps = nreg - 1;
                j.pbase = base;
                j.out = ((int64_t)c * nf + f) * n;
                j.e0off = 0;
                j.pad = 0;
                jobs.push_back(j);
            }
        inverse_and_chains(ctx, n, P.as<zd>(), Pinv.as<zd>(), sel, jobs, steps, de0.as<zd>(), dref.as<zd>(),
                           dboost.as<zd>(), drefl.as<zd>());
        download(reinterpret_cast<zd*>(boost) + (size_t)c0 * nf * n, dboost.p, sizeof(zd) * n * nf * nc, s);
        if (reflect) download(reinterpret_cast<zd*>(refl) + (size_t)c0 * nf * n, drefl.p, sizeof(zd) * n * nf * nc, s);
        BF_HIP(hipStreamSynchronize(s));
    }
}

}  // namespace bf
