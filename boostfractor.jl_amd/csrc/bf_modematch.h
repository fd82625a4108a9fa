// bf_modematch.h -- mode-matching pieces shared by the transformer
// (bf_transformer.hip) and its disk-position sweep caller (bf_sweep.hip).
#pragma once

#include "bf_internal.h"

#include <array>
#include <map>
#include <vector>

namespace bf {

// mode patterns [n][C][ny][nx] and transverse wavenumbers [n] on the device
struct ModesDev {
    int n = 0, C = 0;
    size_t K = 0, Ktot = 0;
    DevBuf pat, kt;
    void upload(const bf_modes* m, size_t K_, hipStream_t s) {
        n = m->n;
        C = m->ncomp;
        K = K_;
        Ktot = (size_t)C * K;
        pat.alloc(sizeof(zd) * n * Ktot);
        kt.alloc(sizeof(zd) * n);
        bf::upload(pat.p, m->patterns, sizeof(zd) * n * Ktot, s);
        bf::upload(kt.p, m->mode_kt, sizeof(zd) * n, s);
    }
};

// Frequency- and geometry-independent pieces of one mode set: the Gram matrix of
// the patterns (flat untilted disks, 1D propagator) and the patterns' FFTs
// (Parseval air regions).  Built on first use, reused by every later geometry.
struct ModeCache {
    DevBuf G, Fhat, FhatM;
    bool haveG = false, haveF = false;
    DevBuf Ginv;  // inv(G), formed by the first flat-disk flush that needs it
    bool haveGinv = false, ginv_ok = false;
    // Parseval regions by k-bin group: bins with equal (kx^2, ky^2) (and, on a
    // square grid with kX == kY, the swapped pair) share one weight, so
    // P = sum_g w_g H_g with H_g = sum_{k in g} conj(Fhat_i(k)) FhatM_j(k).
    bool haveH = false;
    DevBuf H;  // [ngroups][n][n] (column-major)
    // H Hermitian (masked and unmasked pattern transforms equal): its upper triangles packed, Hp[g][t] =
    // H[g][i + n j] for tri[t] = i | j << 16 (i <= j, column by column), nt = n (n + 1) / 2 entries each
    bool herm = false;
    int nt = 0;
    DevBuf Hp, tri;
    // the grouping itself depends on the grid only
    bool haveIdx = false;
    int ngroups = 0;
    DevBuf gkx, gky;    // representative kX, kY of each group (double)
    double gk2max = 0.0;  // max over groups of kX^2 + kY^2
    int gmax = 0;         // most bins in one group
    DevBuf goff, gbin;  // group g holds bins gbin[goff[g] .. goff[g+1])
    // forget everything computed from the patterns (keeps the grouping)
    void reset_data() {
        haveG = haveF = haveH = haveGinv = false;
        reset_series();
    }
    // coefficient matrices of the spacing series (parseval_flush) formed so far in this run, by
    // (k0^2 bits, spacing bin); bin -1 holds dz = 0 for every k0
    struct SeriesEntry {
        const zd* d;  // q coefficient matrices, nn apart
        int q;
        double c, h;
        const zd* e = nullptr;  // q coefficient matrices of the inverse (flat path), nn apart, or none
    };
    std::map<std::array<uint64_t, 3>, SeriesEntry> series;
    // device memory of the coefficients: blocks kept across runs and handed out again from the start by
    // each run (kernels of consecutive runs are ordered on the stream), so a run allocates nothing once
    // the blocks suffice
    struct Arena {
        std::vector<DevBuf> blocks;
        size_t cur = 0, used = 0;
        void reset() { cur = used = 0; }
        void* take(size_t bytes) {
            bytes = (bytes + 255) & ~(size_t)255;
            for (; cur < blocks.size(); cur++, used = 0)
                if (used + bytes <= blocks[cur].bytes) {
                    void* p = static_cast<unsigned char*>(blocks[cur].p) + used;
                    used += bytes;
                    return p;
                }
            blocks.emplace_back(std::max(bytes, (size_t)64 << 20));
            cur = blocks.size() - 1;
            used = bytes;
            return blocks[cur].p;
        }
    } series_mem;
    double series_w = 0.0;  // spacing bin width, fixed by the first flush of the run
    void reset_series() {
        series.clear();
        series_mem.reset();
        series_w = 0.0;
    }
};

// Parseval weight of one (region, frequency): exp(-1im*kz*dz), kz from k0^2
struct PwDesc {
    zd k0sq;
    double dz;
    int pad0, pad1;
};

// Parseval problems of one mode set waiting for one grouped GEMM: their weights
// and the device address of each output matrix
struct ParsevalQueue {
    std::vector<PwDesc> pd;
    std::vector<zd*> dst;
    // flat untilted disks waiting for one k_disk_diag launch: matrix slot (counted
    // in n*n blocks from diag_base), from_gram flag, k0^2 and thickness
    zd* diag_base = nullptr;
    std::vector<int> dslot, dfrom;
    std::vector<zd> dk0sq;
    std::vector<double> ddz;
};

// P = sum_g W_g H_g for every queued problem (GEMMs in chunks bounded by memory); empties the queue
// (with a stager: uploads staged, no wait at the end)
void parseval_flush(bf_ctx* ctx, ModeCache& mc, int n, double scale, ParsevalQueue& q, Stager* st = nullptr);
// the queued disk regions of `q` (diag(exp(-i kz dz)) from the left) in one launch
// inv(G) formed if needed; true if it may stand in for the LU of the flat disks' matrices
bool ginv_usable(bf_ctx* ctx, ModeCache& mc, const ModesDev& md);
// with pinv_base, disks taken from the Gram matrix also get their inverses (from inv(G)) at the same slots
void disk_diag_flush(bf_ctx* ctx, ModeCache& mc, const ModesDev& md, ParsevalQueue& q, Stager* st = nullptr,
                     zd* pinv_base = nullptr);

// one step of the transfer-matrix chain (src/boost_fractor_transformer.jl:326-341)
struct ChainStep {
    zd S;      // axion_S_factor(sqrt(eps[idx(s+1)]), sqrt(eps[idx(s)]))
    zd a, b;   // G = [a I, b I; b I, a I]
    int reg;   // 0-based region index idx_reg(s) - 1
    int disk;  // 1: the region's matrix is diag(ph) G and its inverse inv(G) diag(1/ph) (ChainSrc), else P, Pinv
};

// Where the chain kernels find a region's matrix and inverse: slot m (= job.pbase + step.reg) at P + m n^2
// and Pinv + m n^2, or, for a step marked disk, entry (i, j) as ph_m[i] G[i][j] and Ginv[i][j] phinv_m[j]
// with ph_m = ph + m n, phinv_m = phinv + m n (the flat disks' phases exp(-1im kz dz) and their
// reciprocals; the disks' matrices are never formed)
struct ChainSrc {
    const zd* P;
    const zd* Pinv;
    const zd* G;
    const zd* Ginv;
    const zd* ph;
    const zd* phinv;
};

// One chain per block: the steps [step0, step0 + nsteps), propagation matrices
// P[pbase + reg], the axion source e0 + e0off and the outputs at boost + out
// (and refl + out), n entries each.  A single geometry has one job per
// frequency; a batch of geometries with the same mode count has one job per
// (geometry, frequency), written straight to its place in the results.
struct ChainJob {
    int64_t step0, pbase, out, e0off;
    int nsteps, pad;
};

// propagation_matrix (src/boost_fractor_transformer.jl:174-227) of the regions
// `regs` at nf frequencies into P [nf][nreg][n][n] (column-major matrices).
// `cache` (optional) carries the mode set's frequency-independent pieces across calls.
void propagation_matrices_core(bf_ctx* ctx, const Grid& g, const ModesDev& md, int kind, double diskR, int nreg,
                               const double* distance, const zd* eps, const double* tx, const double* ty,
                               const zd* surfaces_dev, const std::vector<bool>& flat, const double* fr, int nf,
                               const std::vector<int>& regs, zd* P, ModeCache* cache = nullptr,
                               ParsevalQueue* queue = nullptr);

// surface maps on the device and which regions are flat (all-zero map)
void upload_surfaces(bf_ctx* ctx, const bf_setup* st, size_t K, DevBuf& surf, std::vector<bool>& flat);

// permittivities with NaN replaced by 1e30 (src/boost_fractor_transformer.jl:289)
std::vector<zd> eps_fixed(const bf_setup* st);

// chain steps of one geometry at one frequency (regions Nreg-1 .. 1, :311-341)
void append_chain_steps(const std::vector<zd>& eps, int nreg, std::vector<ChainStep>& steps);

// inv(P) of the matrices `sel`, then one chain per job; boost (and refl) [njobs][n]
void inverse_and_chains(bf_ctx* ctx, int n, zd* P, zd* Pinv, const std::vector<int>& sel,
                        const std::vector<ChainJob>& jobs, const std::vector<ChainStep>& steps, const zd* e0,
                        const zd* reflect, zd* boost, zd* refl, Stager* st = nullptr,
                        const ChainSrc* disks = nullptr);

}  // namespace bf

// A resident transformer sweep: inputs uploaded once; every run recomputes
// everything from them (pattern transforms, group Grams, propagation matrices,
// inverses, chains) on the device.  Only the grouping of k bins (a function of
// the grid alone) is kept between runs.
struct bf_transformer_plan {
    bf_ctx* ctx = nullptr;
    bf::Grid g;
    bf::ModesDev md;
    bf::DevBuf surf, P, Pinv, de0, dref, dboost, drefl;
    std::vector<bool> flat;
    std::vector<double> distance, tx, ty, freqs;
    std::vector<zd> eps;
    int nreg = 0, nf = 0, kind = 0;
    double diskR = 0.0;
    bool has_reflect = false;
    bf::ModeCache mc;
};

// Device-resident disk-position sweep: grid, mode sets (patterns), their axion
// sources and the frequencies stay on the device; each run takes the candidate
// geometries from the host and leaves every boost vector in `dboost`.
struct bf_transformer_batch_plan {
    bf_ctx* ctx = nullptr;
    std::vector<bf::Grid> grids;   // one grid for every mode set, or one per mode set
    std::vector<bf::ModesDev> md;  // per mode set
    std::vector<bf::DevBuf> de0;   // per mode set: axion source, complex [n]
    std::vector<bf::ModeCache> mc; // per mode set: the k-bin grouping (grid only) kept across runs
    std::vector<double> freqs;
    int kind = 0;
    double diskR = 0.0;
    bf::DevBuf dboost;  // complex, the last run's results at the caller's offsets
    size_t nboost = 0;  // extent of the last run's results (complex elements)
    std::vector<std::pair<size_t, size_t>> blocks;  // the last run's (offset, length) per geometry
    bf::Stager stager;  // pinned staging of the chunks' descriptor uploads
    bf::DevBuf dbad;    // int: a key's inverse series did not converge in this run
    int64_t reruns = 0; // runs redone with the LU inverses for that reason
    // the grid of mode set `set` (its patterns, and the surface maps of its geometries, are sampled on it)
    const bf::Grid& grid(int set) const { return grids[grids.size() == 1 ? 0 : (size_t)set]; }
};

// cubic B-spline coefficients of a propagation-matrix grid over the spacing axis
struct bf_spacing_itp {
    bf_ctx* ctx = nullptr;
    int nreg = 0, nf = 0, ns = 0, n = 0;
    double first = 0.0, step = 0.0;
    bf::DevBuf coef;  // [ns + 2][nf][nreg][n][n]
};
