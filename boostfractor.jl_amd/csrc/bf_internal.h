// bf_internal.h -- host-side internals shared by the engine translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/boostfractor_hip.h"
#include "bf_device.h"

namespace bf {

// Internal failure carrying a C-ABI status code; converted to a return code at
// the extern "C" boundary.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define BF_HIP(call)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            throw ::bf::Error(BF_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));        \
    } while (0)

#define BF_CHECK_LAUNCH() BF_HIP(hipGetLastError())

// Owning device allocation.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t n) {
        release();
        if (n == 0) return;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            p = nullptr;
            throw Error(e == hipErrorOutOfMemory ? BF_ERR_NOMEM : BF_ERR_HIP,
                        std::string("hipMalloc(") + std::to_string(n) + " bytes): " + hipGetErrorString(e));
        }
        bytes = n;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <typename U>
    U* as() const { return static_cast<U*>(p); }
};

}  // namespace bf

// Launch timing of one kernel class: HIP events recorded on the context's
// stream around each launch, resolved into milliseconds when queried.
struct bf_kernel_timer {
    double ms = 0.0, flops = 0.0;
    int64_t launches = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
};

namespace bf {
// default bytes of P + inv(P) per chunk of a batched sweep: large chunks mean fewer, fuller launches
// (measured on the 4096-geometry sweep: 2 GiB 418k, 4 GiB 450k, 12 GiB 475k, 24 GiB 481k points/s)
constexpr size_t BF_DEFAULT_BATCH_BYTES = (size_t)16 << 30;

// Switches of the engine's work reductions, launch shapes and chunk budgets.
// Each context holds its own set (bf_ctx_set_option); the defaults are the
// product configuration, the other values exist for A/B checks and tuning.
// No setting changes a result beyond the rounding the tests bound.
struct EngineOptions {
    int fft_lds = 0;          // 1: LDS FFT passes instead of the register-resident ones
    int prune = 1;            // band-pruned register transforms
    int qtab = 1;             // quarter k-space tables for complex64 fields
    int band = 1;             // skip the rows the aperture zeroes
    int chfuse = 1;           // cheerleader: pair steps fused with their boundary update
    int chxf = 1;             // cheerleader: pair steps store the next masked row transform
    int chwave = 1;           // cheerleader: wavefront order when amin is 0
    int graph = 1;            // dancer plans replay captured HIP graphs
    int col_tx = 0, rows = 0; // dancer launch shapes (0: tuned/default)
    int parseval_groups = 1;  // flat air regions by k-bin groups
    int disk_expansion = 1;   // tilted/rough disks by a frequency series of moment Grams
    int pg_wn = 0;            // Parseval product: waves along the matrix entries, 2 or 4 (0: by n)
    int parseval_series = 1;  // air gaps of one k0 with nearby spacings by a Taylor series in the spacing
    int series_inverse = 1;   // their inverses by the inverse's series (flat sweep chunks) instead of the LU
    int chain_tile = 1;       // register-tiled chain kernels for n = 9, 15, 25, 45 (0: the LDS chain kernel)
    int chain_mfma = 1;       // chain on the matrix cores: 1 where it pays, 4 for every n <= 32 (products in
                              // separate waves), 2 / 3 for every n <= 32 (both products per wave, four / three
                              // MFMAs per complex product), 0 never
    int chain_order = 1;      // chain blocks start longest chain first (0: in job order)
    int pgemm_herm = 1;       // Parseval product over the upper triangles when the Grams are Hermitian
    int sinv_mfma = 1;        // inverse series on the matrix cores for 8 < n <= 32
    int autotune = 0;         // 1: time candidate launch shapes on first use of a shape key (kept per context)
    int64_t batch_bytes = 0;  // matrix budget of a sweep chunk (0: BF_DEFAULT_BATCH_BYTES)
};
// Options of the context whose call runs on this thread (defaults outside a call).
const EngineOptions& opts();
// Makes `o` the running call's options for the lifetime of the object.
struct OptionScope {
    const EngineOptions* prev;
    explicit OptionScope(const EngineOptions* o);
    ~OptionScope();
    OptionScope(const OptionScope&) = delete;
    OptionScope& operator=(const OptionScope&) = delete;
};
}  // namespace bf

// Per-caller context: one device, one stream, cached twiddle tables.
struct bf_ctx {
    bf::EngineOptions opt;
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // twiddle tables keyed by (n, bytes per real)
    std::map<std::pair<int, int>, bf::DevBuf> twiddles;
    bool timing = false;
    std::map<std::string, bf_kernel_timer> timers;
    // grow-only device buffers reused by consecutive calls on this context
    // (work arrays of the mode-matching batches), released with the context
    std::map<int, bf::DevBuf> scratch;
    // launch shapes picked by the autotuner, by shape key (also written to
    // tune_file when the caller names one), and the time spent tuning
    std::map<std::string, std::vector<int64_t>> tuned;
    std::string tune_file;
    double tune_ms = 0.0;
};

namespace bf {
// a tuned entry was found: append it to the context's tuning file, if any
void save_tuned(bf_ctx* ctx, const std::string& key);

// scratch slot `slot` of the context with at least `bytes` bytes (contents undefined); a buffer that
// grows is freed only after the work queued on the context's stream (which may still use it) is done
inline DevBuf& scratch_buf(bf_ctx* ctx, int slot, size_t bytes) {
    DevBuf& b = ctx->scratch[slot];
    if (b.bytes < bytes) {
        if (b.p) BF_HIP(hipStreamSynchronize(ctx->stream));
        b.alloc(bytes);
    }
    return b;
}

// Host-to-device uploads through pinned staging memory for a sequence of
// chunks whose host preparation overlaps the device work of the previous one:
// put() copies into the current slot's pinned arena and queues an asynchronous
// copy; end_chunk() records an event; begin_chunk() switches to the other slot
// and waits, if needed, until the chunk that last used it has completed.
struct Stager {
    struct Block {
        void* p = nullptr;
        size_t bytes = 0, used = 0;
    };
    struct Slot {
        std::vector<Block> blocks;
        hipEvent_t done = nullptr;
        bool pending = false;
    };
    Slot slot[2];
    int cur = 0;
    Stager() = default;
    Stager(const Stager&) = delete;
    Stager& operator=(const Stager&) = delete;
    ~Stager() {
        for (Slot& sl : slot) {
            if (sl.pending) (void)hipEventSynchronize(sl.done);
            for (Block& b : sl.blocks) (void)hipHostFree(b.p);
            if (sl.done) (void)hipEventDestroy(sl.done);
        }
    }
    void begin_chunk() {
        cur ^= 1;
        Slot& sl = slot[cur];
        if (sl.pending) BF_HIP(hipEventSynchronize(sl.done));
        sl.pending = false;
        if (sl.blocks.size() > 1) {  // merge into one block of the total size
            size_t total = 0;
            for (Block& b : sl.blocks) {
                total += b.bytes;
                BF_HIP(hipHostFree(b.p));
            }
            sl.blocks.clear();
            Block nb;
            BF_HIP(hipHostMalloc(&nb.p, total));
            nb.bytes = total;
            sl.blocks.push_back(nb);
        }
        for (Block& b : sl.blocks) b.used = 0;
    }
    void put(hipStream_t s, void* dst, const void* src, size_t bytes) {
        if (!bytes) return;
        Slot& sl = slot[cur];
        const size_t need = (bytes + 255) & ~(size_t)255;
        if (sl.blocks.empty() || sl.blocks.back().used + need > sl.blocks.back().bytes) {
            Block nb;
            nb.bytes = std::max(need, (size_t)4 << 20);
            BF_HIP(hipHostMalloc(&nb.p, nb.bytes));
            sl.blocks.push_back(nb);
        }
        Block& b = sl.blocks.back();
        void* h = static_cast<char*>(b.p) + b.used;
        b.used += need;
        std::memcpy(h, src, bytes);
        BF_HIP(hipMemcpyAsync(dst, h, bytes, hipMemcpyHostToDevice, s));
    }
    void end_chunk(hipStream_t s) {
        Slot& sl = slot[cur];
        if (!sl.done) BF_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
        BF_HIP(hipEventRecord(sl.done, s));
        sl.pending = true;
    }
};
// upload through the stager when there is one (no wait), else directly
void upload(void* dst, const void* src, size_t bytes, hipStream_t s);
inline void stage_upload(Stager* st, void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (st) st->put(s, dst, src, bytes);
    else upload(dst, src, bytes, s);
}
enum ScratchSlot : int {
    SB_PW = 1, SB_DST, SB_SEL, SB_JOBS, SB_STEPS, SB_TB, SB_CW, SB_P, SB_PINV, SB_BOOST,
    SB_SLOTS, SB_FLAG, SB_K0, SB_DZ, SB_PGSLOT, SB_PGNK, SB_PGK0,
    SB_SKEY, SB_SPROB, SB_SSLICE, SB_SROWDST, SB_SW, SB_SD, SB_PGPART, SB_FENT, SB_FDISK, SB_FK0, SB_FINV, SB_SINVJ, SB_SINVSEL, SB_FISLICE, SB_FPH, SB_HFLAG
};
}  // namespace bf

namespace bf {

// Brackets one launch of kernel class `name` with events while the context's
// timing is on; `flops` is the launch's algorithmic work.  Construct it right
// before the launch, destroy it right after.
struct TimedLaunch {
    bf_ctx* ctx;
    hipEvent_t e1 = nullptr;
    TimedLaunch(bf_ctx* c, const char* name, double flops) : ctx(c) {
        if (!c->timing) return;
        hipEvent_t e0;
        BF_HIP(hipEventCreate(&e0));
        BF_HIP(hipEventCreate(&e1));
        BF_HIP(hipEventRecord(e0, c->stream));
        bf_kernel_timer& t = c->timers[name];
        t.pending.push_back({e0, e1});
        t.flops += flops;
        t.launches += 1;
    }
    ~TimedLaunch() {
        if (e1) (void)hipEventRecord(e1, ctx->stream);
    }
    TimedLaunch(const TimedLaunch&) = delete;
    TimedLaunch& operator=(const TimedLaunch&) = delete;
};
void resolve_timers(bf_ctx* ctx);
void clear_timers(bf_ctx* ctx);

// Device-resident coordinate grids.
struct Grid {
    int nx = 0, ny = 0;
    DevBuf X, Y, kX, kY;  // double
    std::vector<double> hkX, hkY;  // host copies of kX, kY
    void upload(const bf_coords* c, hipStream_t s);
    size_t K() const { return (size_t)nx * ny; }
};

template <typename T>
const cplx<T>* twiddle_table(bf_ctx* ctx, int n);

// per-field descriptor of the propagation passes
struct FieldDesc {
    int64_t off;  // element offset of the field in its buffer
    int ktab;     // k-space table index (-1: none)
    int rtab;     // real-space (tilt*surface) table index (-1: none)
    int group;    // activity group (frequency slot), index into `active`
    int pad;
};

struct KTabDesc {
    zd k0sq;      // k0*k0 (Julia complex product)
    double dz;
    int shifted;  // 1: table laid out in fftshifted order (momentum-space propagator)
    int pad;
};

struct RTabDesc {
    zd a;          // (-1im)*k0
    double tx, ty;
    int64_t surf;  // element offset into the surfaces buffer, -1: flat
};

// allow a kernel more than the default dynamic LDS (gfx950 has 160 KB per CU)
void allow_lds(const void* kernel, size_t bytes);

// helpers used by the mode-matching unit
void fft_fields(bf_ctx* ctx, const Grid& g, zd* base, int nf, bool mask, double diskR);
void propagate_fields(bf_ctx* ctx, const Grid& g, int kind, double diskR, const zd* surfaces, zd* base, int nf,
                      zd eps, double dz, double tx, double ty, int64_t surf_off, double lambda, zd* scratch);
void build_ktabs(bf_ctx* ctx, const Grid& g, const std::vector<KTabDesc>& d, double scale, zd* out);
void build_rtabs(bf_ctx* ctx, const Grid& g, const std::vector<RTabDesc>& d, const zd* surfaces, zd* out);

struct DancerOptions {
    const double* emit = nullptr;     // host [R][2][K]
    const double* reflect = nullptr;  // host [K]
    int returnsum = 1, immediatesum = 1;
    int freq_batch = 0;
    int col_tx = 0, rows = 0;  // launch shapes (0: defaults)
    int band = 1;              // skip the rows the aperture zeroes (0: process every row)
    int graph = 0;             // replay each frequency batch's bounce loop from a captured HIP graph
};

struct DancerPlanBase {
    virtual ~DancerPlanBase() = default;
    virtual void run() = 0;
    virtual void fetch(double* Eout, double* power, int32_t* bounces) = 0;
    virtual void set_timing(bool on) = 0;
    virtual void kernel_times(double* ms, int64_t* launches) = 0;
    bf_ctx* ctx = nullptr;
    int freq_batch = 0;             // frequencies per batch
    int band_lo = 0, band_hi = 0;   // row band of the fused bounce
};

// recursive Fourier propagation (bf_engine.hip)
void propagate_one(bf_ctx* ctx, int kind, const bf_coords* coords, double* E, double dz, double diskR, zd eps,
                   double tx, double ty, const double* surface, double lambda);
void dance_intro_dev(bf_ctx* ctx, const bf_coords* coords, int nd, const double* eps, zd r1, double diskR,
                     double* bfield, double velocity_x, double f, double* out);
void gauss_profile_dev(bf_ctx* ctx, const bf_coords* co, double z, double omega0, double f, double* out);
void unnormalized_match_dev(bf_ctx* ctx, int nb, int64_t n, const double* Eout, const double* Eant, double dx,
                            double dy, double* out);
DancerPlanBase* make_dancer_plan(bf_ctx* ctx, const bf_setup* st, const bf_coords* co, double amin, int nmax,
                                 const double* fr, int nfr, int kind, double diskR, int precision,
                                 const DancerOptions& o);
template <typename T>
void cheerleader_run(bf_ctx* ctx, const bf_setup* st, const bf_coords* co, double amin, int nmax, const double* fr,
                     int nfr, int kind, const double* emit_h, const double* reflect_h, double diskR, double* out_right,
                     double* out_left);

double stream_copy_probe(bf_ctx* ctx, int64_t bytes, int reps);

// mode matching (bf_transformer.hip)
void seed_modes_dev(bf_ctx* ctx, const bf_coords* coords, int M, int L, double diskR, double* patterns, double* kt);
void field2modes_dev(bf_ctx* ctx, const bf_coords* coords, const bf_modes* modes, const double* B, double velocity_x,
                     double diskR, double f, double* coeffs);
void modes2field_dev(bf_ctx* ctx, const bf_coords* coords, const bf_modes* modes, const double* coeffs,
                     double* field);
void propagation_matrices_dev(bf_ctx* ctx, const bf_setup* st, const bf_coords* co, const bf_modes* md,
                              const double* freqs, int nfreq, int kind, double diskR, int precision, double* out);
void transformer_dev(bf_ctx* ctx, const bf_setup* st, const bf_coords* co, const bf_modes* md, const double* freqs,
                     int nfreq, int kind, const double* prop_mats, double diskR, const double* emit,
                     const double* reflect, int precision, double* boost, double* refl);
void disk_system_1d_dev(bf_ctx* ctx, const double* freqs, int nfreq, int ncand, int num_disk, const double* spacings,
                        double thickness, const double* eps_disk, const double* eps_gap, int mirror,
                        const double* surfaceloss, double* gamma, double* axion);
void transform_surface_1d_dev(bf_ctx* ctx, const double* freqs, int n, const double* gamma_in, double* axion,
                              const double* eps_i, const double* eps_m, int m_real, double l, double surfaceloss,
                              int ray, double* gamma_out);
bf_transformer_plan* transformer_plan_create(bf_ctx* ctx, const bf_setup* st, const bf_coords* co,
                                             const bf_modes* modes, const double* freqs, int nfreq, int kind,
                                             double diskR, const double* emit, const double* reflect);
void transformer_plan_run(bf_transformer_plan* pl);
void transformer_plan_fetch(bf_transformer_plan* pl, double* boost, double* refl);
void transformer_plan_destroy(bf_transformer_plan* pl);
void transformer_batch_dev(bf_ctx* ctx, int nconf, const bf_setup* setups, const bf_coords* co, int nsets,
                           const bf_modes* modesets, const double* const* emits, const int32_t* set_of,
                           const double* freqs, int nfreq, int kind, double diskR, const int64_t* boost_off,
                           double* boost);
// co: one grid (ncoords = 1) or one per mode set (ncoords = nsets)
bf_transformer_batch_plan* transformer_batch_plan_create(bf_ctx* ctx, const bf_coords* co, int ncoords, int nsets,
                                                         const bf_modes* modesets, const double* const* emits,
                                                         const double* freqs, int nfreq, int kind, double diskR);
void transformer_batch_plan_run(bf_transformer_batch_plan* pl, int nconf, const bf_setup* setups,
                                const int32_t* set_of, const int64_t* boost_off);
void transformer_batch_plan_fetch(bf_transformer_batch_plan* pl, double* boost);
void transformer_batch_plan_destroy(bf_transformer_batch_plan* pl);
void propagation_matrices_grid_dev(bf_ctx* ctx, const bf_setup* st, const bf_coords* co, const bf_modes* modes,
                                   const double* spacing, int ns, const double* tgx, int ntx, const double* tgy,
                                   int nty, const double* freqs, int nf, int kind, double diskR, double* out);
bf_spacing_itp* spacing_itp_create(bf_ctx* ctx, const double* grid, int nreg, int nf, int ns, int n, double first,
                                   double step);
void spacing_itp_destroy(bf_spacing_itp* it);
void spacing_itp_bounds(const bf_spacing_itp* it, double* lo, double* hi);
void spacing_itp_matrices(bf_spacing_itp* it, int ncand, const double* shifts, double* out);
void spacing_itp_transformer(bf_spacing_itp* it, int ncand, const double* shifts, const double* eps,
                             const double* emit, const double* reflect, double* boost, double* refl);
constexpr int BF_NOALLOC_MAX_DEVICES = 64;
// device address of the allocation-free entry point's scratch array on `device`
// (taken once per device, with the device current); bytes a call needs
void noalloc_init(int device);
size_t noalloc_bytes(int nreg, int n, int nfreq);
void transformer_chain_noalloc(bf_ctx* ctx, int nreg, const double* eps, int n, int nfreq, const zd* P, const zd* e0,
                               const zd* reflect, zd* boost, zd* refl);
void trace_back_dev(bf_ctx* ctx, const double* reflected, const double* input, const bf_setup* st,
                    const bf_coords* co, const bf_modes* md, double f, int kind, const double* prop_mats,
                    double diskR, int includes_axion, const double* emit, double* fields);

void upload(void* dst, const void* src, size_t bytes, hipStream_t s);
void download(void* dst, const void* src, size_t bytes, hipStream_t s);

}  // namespace bf
