// bf_capi.hip -- extern "C" entry points of libboostfractor_hip.so.
//
// Every entry point validates its arguments on the host first (so the error
// paths are testable without a device), then runs the engine, converting any
// internal bf::Error into a status code and a message for bf_last_error().
#include "bf_modematch.h"

#include <cmath>
#include <cstdio>
#include <cstring>


struct bf_dancer_plan {
    bf::DancerPlanBase* impl = nullptr;
};

static thread_local std::string g_last_error;

static int fail(const bf_ctx* ctx, int code, const std::string& msg) {
    g_last_error = msg;
    if (ctx) const_cast<bf_ctx*>(ctx)->err = msg;
    return code;
}

#define BF_REQUIRE(ctx, cond, code, msg) \
    do {                                 \
        if (!(cond)) return fail(ctx, code, msg); \
    } while (0)

namespace bf {
static thread_local const EngineOptions* t_opt = nullptr;
const EngineOptions& opts() {
    static const EngineOptions defaults;
    return t_opt ? *t_opt : defaults;
}
OptionScope::OptionScope(const EngineOptions* o) : prev(t_opt) { t_opt = o; }
OptionScope::~OptionScope() { t_opt = prev; }
}  // namespace bf

namespace bf {
// one line per entry: the key, a tab, the values separated by spaces
void save_tuned(bf_ctx* ctx, const std::string& key) {
    if (ctx->tune_file.empty()) return;
    FILE* f = std::fopen(ctx->tune_file.c_str(), "a");
    if (!f) return;  // a read-only location only loses the cache
    std::fprintf(f, "%s\t", key.c_str());
    for (int64_t v : ctx->tuned[key]) std::fprintf(f, " %lld", (long long)v);
    std::fprintf(f, "\n");
    std::fclose(f);
}
}  // namespace bf

template <typename F>
static int guarded(bf_ctx* ctx, F&& body) {
    try {
        bf::OptionScope scope(ctx ? &ctx->opt : nullptr);
        if (ctx) BF_HIP(hipSetDevice(ctx->device));
        body();
        return BF_OK;
    } catch (const bf::Error& e) {
        return fail(ctx, e.code, e.what());
    } catch (const std::bad_alloc&) {
        return fail(ctx, BF_ERR_NOMEM, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(ctx, BF_ERR_HIP, e.what());
    }
}

// ---- argument checks ------------------------------------------------------
static int check_coords(const bf_ctx* ctx, const bf_coords* c) {
    BF_REQUIRE(ctx, c != nullptr, BF_ERR_ARG, "coords is NULL");
    BF_REQUIRE(ctx, c->nx >= 1 && c->ny >= 1, BF_ERR_DIM, "grid dimensions must be >= 1");
    BF_REQUIRE(ctx, c->nx <= BF_MAX_GRID && c->ny <= BF_MAX_GRID, BF_ERR_UNSUPPORTED,
               "grid dimensions above 1024 are not supported");
    BF_REQUIRE(ctx, c->X && c->Y && c->kX && c->kY, BF_ERR_ARG, "coordinate arrays must not be NULL");
    return BF_OK;
}

static int check_setup(const bf_ctx* ctx, const bf_setup* s, int min_reg, bool need_r) {
    BF_REQUIRE(ctx, s != nullptr, BF_ERR_ARG, "setup is NULL");
    BF_REQUIRE(ctx, s->nreg >= min_reg, BF_ERR_DIM,
               "the arrays in your SetupBoundaries objects don't fit together! (too few regions)");
    BF_REQUIRE(ctx, s->distance && s->eps && s->tilt_x && s->tilt_y, BF_ERR_ARG, "setup arrays must not be NULL");
    BF_REQUIRE(ctx, !need_r || s->r, BF_ERR_ARG, "setup reflectivities must not be NULL");
    return BF_OK;
}

static int check_kind(const bf_ctx* ctx, int kind) {
    BF_REQUIRE(ctx, kind >= BF_PROP_PROPAGATOR && kind <= BF_PROP_MOMENTUM, BF_ERR_ARG,
               "unknown propagator kind (user-defined Julia propagators have no device equivalent)");
    return BF_OK;
}

static int check_prec(const bf_ctx* ctx, int p) {
    BF_REQUIRE(ctx, p == BF_PREC_FP64 || p == BF_PREC_MIXED || p == BF_PREC_FP32, BF_ERR_ARG, "unknown precision");
    return BF_OK;
}

static int check_freqs(const bf_ctx* ctx, const double* f, int n) {
    BF_REQUIRE(ctx, f != nullptr && n >= 1, BF_ERR_ARG, "at least one frequency is required");
    for (int i = 0; i < n; i++) BF_REQUIRE(ctx, std::isfinite(f[i]) && f[i] > 0.0, BF_ERR_ARG, "frequencies must be > 0");
    return BF_OK;
}

static int check_modes(const bf_ctx* ctx, const bf_modes* m) {
    BF_REQUIRE(ctx, m != nullptr && m->patterns && m->mode_kt, BF_ERR_ARG, "modes or its arrays are NULL");
    BF_REQUIRE(ctx, m->n >= 1 && m->ncomp >= 1, BF_ERR_DIM, "modes must have n >= 1 and ncomp >= 1");
    BF_REQUIRE(ctx, m->n <= BF_MAX_MODES, BF_ERR_UNSUPPORTED, "more than 128 modes are not supported");
    return BF_OK;
}

#define TRY(x)              \
    do {                    \
        int rc_ = (x);      \
        if (rc_) return rc_; \
    } while (0)

extern "C" {

int bf_version(void) { return 10000; }

const char* bf_last_error(const bf_ctx* ctx) {
    if (ctx && !ctx->err.empty()) return ctx->err.c_str();
    return g_last_error.c_str();
}

int bf_device_count(int32_t* count) {
    if (!count) return fail(nullptr, BF_ERR_ARG, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        *count = 0;
        return fail(nullptr, BF_ERR_HIP, std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
    }
    *count = n;
    return BF_OK;
}

int bf_ctx_create(int32_t device, bf_ctx** out) {
    if (!out) return fail(nullptr, BF_ERR_ARG, "out is NULL");
    *out = nullptr;
    bf_ctx* c = new bf_ctx();
    c->device = device;
    int rc = guarded(nullptr, [&] {
        int n = 0;
        BF_HIP(hipGetDeviceCount(&n));
        if (device < 0 || device >= n) throw bf::Error(BF_ERR_ARG, "device index out of range");
        BF_HIP(hipSetDevice(device));
        BF_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        bf::noalloc_init(device);  // the allocation-free entry point's scratch address on this device
    });
    if (rc) {
        delete c;
        return rc;
    }
    *out = c;
    return BF_OK;
}

int bf_ctx_destroy(bf_ctx* ctx) {
    if (!ctx) return BF_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    bf::clear_timers(ctx);
    ctx->twiddles.clear();
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return BF_OK;
}

int bf_ctx_stream(bf_ctx* ctx, void** stream) {
    BF_REQUIRE(ctx, ctx && stream, BF_ERR_ARG, "ctx or stream is NULL");
    *stream = (void*)ctx->stream;
    return BF_OK;
}

int bf_ctx_timing(bf_ctx* ctx, int32_t enable) {
    BF_REQUIRE(ctx, ctx, BF_ERR_ARG, "ctx is NULL");
    return guarded(ctx, [&] {
        bf::clear_timers(ctx);
        ctx->timing = enable != 0;
    });
}

int bf_ctx_kernel_time(bf_ctx* ctx, const char* kernel, double* ms, int64_t* launches, double* flops) {
    BF_REQUIRE(ctx, ctx && kernel, BF_ERR_ARG, "ctx or kernel is NULL");
    return guarded(ctx, [&] {
        bf::resolve_timers(ctx);
        auto it = ctx->timers.find(kernel);
        const bool have = it != ctx->timers.end();
        if (ms) *ms = have ? it->second.ms : 0.0;
        if (launches) *launches = have ? it->second.launches : 0;
        if (flops) *flops = have ? it->second.flops : 0.0;
    });
}

// name -> option field of bf::EngineOptions
static int64_t* option_slot(bf::EngineOptions& o, const char* name, int64_t& tmp, int*& ip) {
    struct Int { const char* n; int bf::EngineOptions::*m; };
    static const Int ints[] = {
        {"fft_lds", &bf::EngineOptions::fft_lds}, {"prune", &bf::EngineOptions::prune},
        {"qtab", &bf::EngineOptions::qtab}, {"band", &bf::EngineOptions::band},
        {"chfuse", &bf::EngineOptions::chfuse}, {"chxf", &bf::EngineOptions::chxf},
        {"chwave", &bf::EngineOptions::chwave}, {"graph", &bf::EngineOptions::graph},
        {"col_tx", &bf::EngineOptions::col_tx}, {"rows", &bf::EngineOptions::rows},
        {"parseval_groups", &bf::EngineOptions::parseval_groups},
        {"disk_expansion", &bf::EngineOptions::disk_expansion}, {"pg_wn", &bf::EngineOptions::pg_wn},
        {"autotune", &bf::EngineOptions::autotune}, {"parseval_series", &bf::EngineOptions::parseval_series},
        {"series_inverse", &bf::EngineOptions::series_inverse},
        {"chain_tile", &bf::EngineOptions::chain_tile},
        {"chain_mfma", &bf::EngineOptions::chain_mfma},
        {"chain_order", &bf::EngineOptions::chain_order},
        {"pgemm_herm", &bf::EngineOptions::pgemm_herm},
        {"sinv_mfma", &bf::EngineOptions::sinv_mfma},
    };
    ip = nullptr;
    for (const Int& e : ints)
        if (!strcmp(name, e.n)) {
            ip = &(o.*e.m);
            tmp = *ip;
            return &tmp;
        }
    if (!strcmp(name, "batch_bytes")) return &o.batch_bytes;
    return nullptr;
}

int bf_ctx_set_option(bf_ctx* ctx, const char* name, int64_t value) {
    BF_REQUIRE(ctx, ctx && name, BF_ERR_ARG, "ctx or name is NULL");
    int64_t tmp = 0;
    int* ip = nullptr;
    int64_t* slot = option_slot(ctx->opt, name, tmp, ip);
    BF_REQUIRE(ctx, slot != nullptr, BF_ERR_ARG, std::string("unknown option: ") + name);
    BF_REQUIRE(ctx, value >= 0 && (!ip || value <= 1 << 30), BF_ERR_ARG, "option values must be >= 0");
    if (ip) *ip = (int)value;
    else *slot = value;
    return BF_OK;
}

int bf_ctx_get_option(bf_ctx* ctx, const char* name, int64_t* value) {
    BF_REQUIRE(ctx, ctx && name && value, BF_ERR_ARG, "ctx, name or value is NULL");
    int64_t tmp = 0;
    int* ip = nullptr;
    int64_t* slot = option_slot(ctx->opt, name, tmp, ip);
    BF_REQUIRE(ctx, slot != nullptr, BF_ERR_ARG, std::string("unknown option: ") + name);
    *value = *slot;
    return BF_OK;
}

int bf_ctx_set_tuning_file(bf_ctx* ctx, const char* path) {
    BF_REQUIRE(ctx, ctx, BF_ERR_ARG, "ctx is NULL");
    ctx->tune_file = path ? path : "";
    if (!path) return BF_OK;
    FILE* f = std::fopen(path, "r");
    if (!f) return BF_OK;  // a new cache
    char line[1024];
    while (std::fgets(line, sizeof line, f)) {
        char* tab = std::strchr(line, '\t');
        if (!tab) continue;
        *tab = 0;
        std::vector<int64_t> v;
        char* p = tab + 1;
        char* end = nullptr;
        for (long long x = std::strtoll(p, &end, 10); end != p; x = std::strtoll(p, &end, 10)) {
            v.push_back(x);
            p = end;
        }
        if (!v.empty()) ctx->tuned[line] = v;  // a later line for a key replaces an earlier one
    }
    std::fclose(f);
    return BF_OK;
}

int bf_ctx_tuning_report(bf_ctx* ctx, char* buf, int64_t len, double* tune_ms) {
    BF_REQUIRE(ctx, ctx, BF_ERR_ARG, "ctx is NULL");
    std::string r;
    for (const auto& e : ctx->tuned) {
        r += e.first + "\t";
        for (int64_t v : e.second) r += " " + std::to_string(v);
        r += "\n";
    }
    if (tune_ms) *tune_ms = ctx->tune_ms;
    if (buf && len > 0) {
        const size_t n = std::min(r.size(), (size_t)len - 1);
        std::memcpy(buf, r.data(), n);
        buf[n] = 0;
    }
    BF_REQUIRE(ctx, !buf || (int64_t)r.size() < len, BF_ERR_ARG, "report buffer too small");
    return BF_OK;
}

int bf_ctx_synchronize(bf_ctx* ctx) {
    BF_REQUIRE(ctx, ctx, BF_ERR_ARG, "ctx is NULL");
    return guarded(ctx, [&] { BF_HIP(hipStreamSynchronize(ctx->stream)); });
}

int bf_propagate(bf_ctx* ctx, int32_t prop_kind, const bf_coords* coords, double* E, double dz, double diskR,
                 double eps_re, double eps_im, double tilt_x, double tilt_y, const double* surface, double lambda) {
    TRY(check_kind(ctx, prop_kind));
    TRY(check_coords(ctx, coords));
    BF_REQUIRE(ctx, E != nullptr, BF_ERR_ARG, "E is NULL");
    BF_REQUIRE(ctx, lambda > 0.0, BF_ERR_ARG, "lambda must be > 0");
    BF_REQUIRE(ctx, ctx != nullptr, BF_ERR_ARG, "ctx is NULL");
    return guarded(ctx, [&] {
        bf::propagate_one(ctx, prop_kind, coords, E, dz, diskR, cmk<double>(eps_re, eps_im), tilt_x, tilt_y, surface,
                          lambda);
    });
}

int bf_dance_intro(bf_ctx* ctx, const bf_coords* coords, int32_t nd, const double* eps, double r1_re, double r1_im,
                   double diskR, double* bfield, double velocity_x, double f, double* out) {
    TRY(check_coords(ctx, coords));
    BF_REQUIRE(ctx, nd >= 1 && eps && out, BF_ERR_ARG, "dance_intro needs nd >= 1, eps and out");
    BF_REQUIRE(ctx, velocity_x == 0.0 || bfield != nullptr, BF_ERR_ARG,
               "InexactError: velocity_x needs a complex bfield (the default B field is real)");
    BF_REQUIRE(ctx, velocity_x == 0.0 || f > 0.0, BF_ERR_ARG, "frequency must be positive");
    BF_REQUIRE(ctx, ctx != nullptr, BF_ERR_ARG, "ctx is NULL");
    return guarded(ctx, [&] {
        bf::dance_intro_dev(ctx, coords, nd, eps, cmk<double>(r1_re, r1_im), diskR, bfield, velocity_x, f, out);
    });
}

int bf_gauss_profile(bf_ctx* ctx, const bf_coords* coords, double z, double omega0, double f, double* out) {
    TRY(check_coords(ctx, coords));
    BF_REQUIRE(ctx, out != nullptr, BF_ERR_ARG, "out is NULL");
    BF_REQUIRE(ctx, omega0 > 0.0 && f > 0.0, BF_ERR_ARG, "gauss_profile needs omega0 > 0 and f > 0");
    BF_REQUIRE(ctx, ctx != nullptr, BF_ERR_ARG, "ctx is NULL");
    return guarded(ctx, [&] { bf::gauss_profile_dev(ctx, coords, z, omega0, f, out); });
}

int bf_unnormalized_match(bf_ctx* ctx, int32_t nfield, int64_t n, const double* Eout, const double* Eantenna,
                          double dx, double dy, double* out) {
    BF_REQUIRE(ctx, nfield >= 1 && n >= 1 && Eout && Eantenna && out, BF_ERR_ARG,
               "unnormalized_match needs nfield >= 1, n >= 1 and non-NULL buffers");
    BF_REQUIRE(ctx, ctx != nullptr, BF_ERR_ARG, "ctx is NULL");
    return guarded(ctx, [&] { bf::unnormalized_match_dev(ctx, nfield, n, Eout, Eantenna, dx, dy, out); });
}

// Launch shapes and the row band from the context's options (bf_ctx_set_option).
static bf::DancerOptions ctx_options() {
    const bf::EngineOptions& e = bf::opts();
    bf::DancerOptions o;
    o.col_tx = e.col_tx;
    o.rows = e.rows;
    o.band = e.band;
    return o;
}

int bf_dancer(bf_ctx* ctx, const bf_setup* setup, const bf_coords* coords, double amin, int32_t nmax,
              const double* freqs, int32_t nfreq, int32_t prop_kind, const double* emit, const double* reflect,
              double diskR, int32_t returnsum, int32_t immediatesum, int32_t precision, double* Eout, double* power,
              int32_t* bounces) {
    TRY(check_kind(ctx, prop_kind));
    TRY(check_prec(ctx, precision));
    TRY(check_coords(ctx, coords));
    TRY(check_setup(ctx, setup, 2, true));
    TRY(check_freqs(ctx, freqs, nfreq));
    BF_REQUIRE(ctx, nmax >= -1, BF_ERR_ARG, "nmax must be >= -1");
    BF_REQUIRE(ctx, Eout || power, BF_ERR_ARG, "Eout and power are both NULL");
    BF_REQUIRE(ctx, ctx != nullptr, BF_ERR_ARG, "ctx is NULL");
    return guarded(ctx, [&] {
        bf::DancerOptions o = ctx_options();
        o.reflect = reflect;  // reflect takes precedence over emit (src/boost_fractor_recurprop.jl:53-63)
        o.emit = reflect ? nullptr : emit;
        o.returnsum = returnsum ? 1 : 0;
        o.immediatesum = immediatesum ? 1 : 0;
        std::unique_ptr<bf::DancerPlanBase> p(bf::make_dancer_plan(ctx, setup, coords, amin, nmax, freqs, nfreq,
                                                                   prop_kind, diskR, precision, o));
        p->run();
        p->fetch(Eout, power, bounces);
    });
}

int bf_dancer_batch(bf_ctx* ctx, int32_t nconf, const bf_setup* setups, const bf_coords* coords, double amin,
                    int32_t nmax, const double* freqs, int32_t nfreq, int32_t prop_kind,
                    const double* const* emits, const double* const* reflects, double diskR, int32_t returnsum,
                    int32_t immediatesum, int32_t precision, const int64_t* eout_offset, double* Eout,
                    double* power, int32_t* bounces) {
    TRY(check_kind(ctx, prop_kind));
    TRY(check_prec(ctx, precision));
    TRY(check_coords(ctx, coords));
    TRY(check_freqs(ctx, freqs, nfreq));
    BF_REQUIRE(ctx, nconf >= 0, BF_ERR_ARG, "nconf must be >= 0");
    BF_REQUIRE(ctx, nconf == 0 || setups, BF_ERR_ARG, "setups is NULL");
    BF_REQUIRE(ctx, nmax >= -1, BF_ERR_ARG, "nmax must be >= -1");
    BF_REQUIRE(ctx, Eout || power, BF_ERR_ARG, "Eout and power are both NULL");
    BF_REQUIRE(ctx, !Eout || eout_offset, BF_ERR_ARG, "Eout needs eout_offset");
    for (int32_t c = 0; c < nconf; c++) {
        TRY(check_setup(ctx, &setups[c], 2, true));
        BF_REQUIRE(ctx, !Eout || eout_offset[c] >= 0, BF_ERR_ARG, "eout_offset must be >= 0");
    }
    BF_REQUIRE(ctx, ctx != nullptr, BF_ERR_ARG, "ctx is NULL");
    return guarded(ctx, [&] {
        // one plan per geometry on the context's stream: geometries may differ in region count
        for (int32_t c = 0; c < nconf; c++) {
            bf::DancerOptions o = ctx_options();
            const double* rf = reflects ? reflects[c] : nullptr;
            o.reflect = rf;  // reflect takes precedence over emit (src/boost_fractor_recurprop.jl:53-63)
            o.emit = rf ? nullptr : (emits ? emits[c] : nullptr);
            o.returnsum = returnsum ? 1 : 0;
            o.immediatesum = immediatesum ? 1 : 0;
            std::unique_ptr<bf::DancerPlanBase> p(bf::make_dancer_plan(ctx, &setups[c], coords, amin, nmax, freqs,
                                                                       nfreq, prop_kind, diskR, precision, o));
            p->run();
            p->fetch(Eout ? Eout + 2 * eout_offset[c] : nullptr, power ? power + (size_t)c * nfreq : nullptr,
                     bounces ? bounces + (size_t)c * nfreq : nullptr);
        }
    });
}

int bf_cheerleader(bf_ctx* ctx, const bf_setup* setup, const bf_coords* coords, double amin, int32_t nmax,
                   const double* freqs, int32_t nfreq, int32_t prop_kind, const double* emit, const double* reflect,
                   double diskR, int32_t precision, double* out_right, double* out_left) {
    TRY(check_kind(ctx, prop_kind));
    TRY(check_prec(ctx, precision));
    TRY(check_coords(ctx, coords));
    TRY(check_setup(ctx, setup, 3, true));
    TRY(check_freqs(ctx, freqs, nfreq));
    BF_REQUIRE(ctx, nmax >= -1, BF_ERR_ARG, "nmax must be >= -1");
    BF_REQUIRE(ctx, out_right != nullptr, BF_ERR_ARG, "out_right is NULL");
    BF_REQUIRE(ctx, ctx != nullptr, BF_ERR_ARG, "ctx is NULL");
    return guarded(ctx, [&] {
        const double* em = reflect ? nullptr : emit;
        if (precision == BF_PREC_MIXED)
            bf::cheerleader_run<float>(ctx, setup, coords, amin, nmax, freqs, nfreq, prop_kind, em, reflect, diskR,
                                       out_right, out_left);
        else if (precision == BF_PREC_FP32)
            bf::cheerleader_run<bf::f32pure>(ctx, setup, coords, amin, nmax, freqs, nfreq, prop_kind, em, reflect,
                                             diskR, out_right, out_left);
        else
            bf::cheerleader_run<double>(ctx, setup, coords, amin, nmax, freqs, nfreq, prop_kind, em, reflect, diskR,
                                        out_right, out_left);
    });
}

int bf_dancer_plan_create(bf_ctx* ctx, const bf_setup* setup, const bf_coords* coords, double amin, int32_t nmax,
                          const double* freqs, int32_t nfreq, int32_t prop_kind, double diskR, int32_t precision,
                          int32_t freq_batch, bf_dancer_plan** out) {
    BF_REQUIRE(ctx, out != nullptr, BF_ERR_ARG, "out is NULL");
    *out = nullptr;
    TRY(check_kind(ctx, prop_kind));
    TRY(check_prec(ctx, precision));
    TRY(check_coords(ctx, coords));
    TRY(check_setup(ctx, setup, 2, true));
    TRY(check_freqs(ctx, freqs, nfreq));
    BF_REQUIRE(ctx, nmax >= -1, BF_ERR_ARG, "nmax must be >= -1");
    BF_REQUIRE(ctx, freq_batch >= 0, BF_ERR_ARG, "freq_batch must be >= 0");
    BF_REQUIRE(ctx, ctx != nullptr, BF_ERR_ARG, "ctx is NULL");
    return guarded(ctx, [&] {
        bf::DancerOptions o = ctx_options();
        o.freq_batch = freq_batch;
        o.graph = bf::opts().graph ? 1 : 0;  // a plan is run many times
        auto* p = new bf_dancer_plan();
        try {
            p->impl = bf::make_dancer_plan(ctx, setup, coords, amin, nmax, freqs, nfreq, prop_kind, diskR, precision, o);
        } catch (...) {
            delete p;
            throw;
        }
        *out = p;
    });
}

int bf_dancer_plan_run(bf_dancer_plan* plan) {
    BF_REQUIRE(nullptr, plan && plan->impl, BF_ERR_ARG, "plan is NULL");
    bf_ctx* ctx = plan->impl->ctx;
    return guarded(ctx, [&] { plan->impl->run(); });
}

int bf_dancer_plan_fetch(bf_dancer_plan* plan, double* Eout, double* power) {
    BF_REQUIRE(nullptr, plan && plan->impl, BF_ERR_ARG, "plan is NULL");
    bf_ctx* ctx = plan->impl->ctx;
    return guarded(ctx, [&] { plan->impl->fetch(Eout, power, nullptr); });
}

int bf_dancer_plan_timing(bf_dancer_plan* plan, int32_t enable) {
    BF_REQUIRE(nullptr, plan && plan->impl, BF_ERR_ARG, "plan is NULL");
    bf_ctx* ctx = plan->impl->ctx;
    return guarded(ctx, [&] { plan->impl->set_timing(enable != 0); });
}

int bf_dancer_plan_kernel_times(bf_dancer_plan* plan, double* ms, int64_t* launches) {
    BF_REQUIRE(nullptr, plan && plan->impl, BF_ERR_ARG, "plan is NULL");
    bf_ctx* ctx = plan->impl->ctx;
    return guarded(ctx, [&] { plan->impl->kernel_times(ms, launches); });
}

int bf_dancer_plan_info(bf_dancer_plan* plan, int32_t* freq_batch, int32_t* band_lo, int32_t* band_hi) {
    BF_REQUIRE(nullptr, plan && plan->impl, BF_ERR_ARG, "plan is NULL");
    const bf::DancerPlanBase* p = plan->impl;
    if (freq_batch) *freq_batch = p->freq_batch;
    if (band_lo) *band_lo = p->band_lo;
    if (band_hi) *band_hi = p->band_hi;
    return BF_OK;
}

int bf_dancer_plan_destroy(bf_dancer_plan* plan) {
    if (!plan) return BF_OK;
    if (plan->impl) {
        (void)hipSetDevice(plan->impl->ctx->device);
        (void)hipStreamSynchronize(plan->impl->ctx->stream);
        delete plan->impl;
    }
    delete plan;
    return BF_OK;
}

int bf_stream_copy_probe(bf_ctx* ctx, int64_t bytes, int32_t reps, double* ms_per_copy) {
    BF_REQUIRE(ctx, bytes >= 16 && reps >= 1 && ms_per_copy, BF_ERR_ARG, "bytes >= 16, reps >= 1 and an output");
    BF_REQUIRE(ctx, ctx != nullptr, BF_ERR_ARG, "ctx is NULL");
    return guarded(ctx, [&] { *ms_per_copy = bf::stream_copy_probe(ctx, bytes, reps); });
}

int bf_seed_modes(bf_ctx* ctx, const bf_coords* coords, int32_t M, int32_t L, double diskR, double* patterns,
                  double* mode_kt) {
    TRY(check_coords(ctx, coords));
    BF_REQUIRE(ctx, M >= 1 && L >= 0, BF_ERR_ARG, "SeedModes needs Mmax >= 1 and Lmax >= 0");
    BF_REQUIRE(ctx, (int64_t)M * (2 * L + 1) <= BF_MAX_MODES, BF_ERR_UNSUPPORTED, "more than 128 modes");
    BF_REQUIRE(ctx, patterns && mode_kt, BF_ERR_ARG, "output arrays are NULL");
    BF_REQUIRE(ctx, diskR > 0.0, BF_ERR_ARG, "diskR must be > 0");
    BF_REQUIRE(ctx, ctx != nullptr, BF_ERR_ARG, "ctx is NULL");
    return guarded(ctx, [&] { bf::seed_modes_dev(ctx, coords, M, L, diskR, patterns, mode_kt); });
}

int bf_field2modes(bf_ctx* ctx, const bf_coords* coords, const bf_modes* modes, const double* B, double velocity_x,
                   double diskR, double f, double* coeffs) {
    TRY(check_coords(ctx, coords));
    TRY(check_modes(ctx, modes));
    BF_REQUIRE(ctx, coeffs != nullptr, BF_ERR_ARG, "coeffs is NULL");
    BF_REQUIRE(ctx, velocity_x == 0.0 || f > 0.0, BF_ERR_ARG, "frequency must be > 0");
    BF_REQUIRE(ctx, ctx != nullptr, BF_ERR_ARG, "ctx is NULL");
    return guarded(ctx, [&] { bf::field2modes_dev(ctx, coords, modes, B, velocity_x, diskR, f, coeffs); });
}

int bf_modes2field(bf_ctx* ctx, const bf_coords* coords, const bf_modes* modes, const double* coeffs,
                   double* field) {
    TRY(check_coords(ctx, coords));
    TRY(check_modes(ctx, modes));
    BF_REQUIRE(ctx, coeffs && field, BF_ERR_ARG, "coeffs or field is NULL");
    BF_REQUIRE(ctx, ctx != nullptr, BF_ERR_ARG, "ctx is NULL");
    return guarded(ctx, [&] { bf::modes2field_dev(ctx, coords, modes, coeffs, field); });
}

int bf_propagation_matrices(bf_ctx* ctx, const bf_setup* setup, const bf_coords* coords, const bf_modes* modes,
                            const double* freqs, int32_t nfreq, int32_t prop_kind, double diskR, int32_t precision,
                            double* out) {
    TRY(check_kind(ctx, prop_kind));
    TRY(check_prec(ctx, precision));
    TRY(check_coords(ctx, coords));
    TRY(check_setup(ctx, setup, 1, false));
    TRY(check_modes(ctx, modes));
    TRY(check_freqs(ctx, freqs, nfreq));
    BF_REQUIRE(ctx, out != nullptr, BF_ERR_ARG, "out is NULL");
    BF_REQUIRE(ctx, ctx != nullptr, BF_ERR_ARG, "ctx is NULL");
    return guarded(ctx, [&] {
        bf::propagation_matrices_dev(ctx, setup, coords, modes, freqs, nfreq, prop_kind, diskR, precision, out);
    });
}

int bf_transformer(bf_ctx* ctx, const bf_setup* setup, const bf_coords* coords, const bf_modes* modes,
                   const double* freqs, int32_t nfreq, int32_t prop_kind, const double* prop_mats, double diskR,
                   const double* emit, const double* reflect, int32_t precision, double* boost, double* refl) {
    TRY(check_kind(ctx, prop_kind));
    TRY(check_prec(ctx, precision));
    TRY(check_coords(ctx, coords));
    TRY(check_setup(ctx, setup, 2, false));
    TRY(check_modes(ctx, modes));
    TRY(check_freqs(ctx, freqs, nfreq));
    BF_REQUIRE(ctx, emit && boost, BF_ERR_ARG, "emit and boost must not be NULL");
    BF_REQUIRE(ctx, !reflect || refl, BF_ERR_ARG, "refl output is NULL while reflect is given");
    BF_REQUIRE(ctx, ctx != nullptr, BF_ERR_ARG, "ctx is NULL");
    return guarded(ctx, [&] {
        bf::transformer_dev(ctx, setup, coords, modes, freqs, nfreq, prop_kind, prop_mats, diskR, emit, reflect,
                            precision, boost, refl);
    });
}

int bf_transformer_plan_create(bf_ctx* ctx, const bf_setup* setup, const bf_coords* coords, const bf_modes* modes,
                               const double* freqs, int32_t nfreq, int32_t prop_kind, double diskR, const double* emit,
                               const double* reflect, bf_transformer_plan** out) {
    BF_REQUIRE(ctx, out != nullptr, BF_ERR_ARG, "out is NULL");
    *out = nullptr;
    TRY(check_kind(ctx, prop_kind));
    TRY(check_coords(ctx, coords));
    TRY(check_setup(ctx, setup, 2, false));
    TRY(check_modes(ctx, modes));
    TRY(check_freqs(ctx, freqs, nfreq));
    BF_REQUIRE(ctx, emit != nullptr, BF_ERR_ARG, "emit is NULL");
    BF_REQUIRE(ctx, ctx != nullptr, BF_ERR_ARG, "ctx is NULL");
    return guarded(ctx, [&] {
        *out = bf::transformer_plan_create(ctx, setup, coords, modes, freqs, nfreq, prop_kind, diskR, emit, reflect);
    });
}

int bf_transformer_plan_run(bf_transformer_plan* plan) {
    BF_REQUIRE(nullptr, plan != nullptr, BF_ERR_ARG, "plan is NULL");
    return guarded(plan->ctx, [&] { bf::transformer_plan_run(plan); });
}

int bf_transformer_plan_fetch(bf_transformer_plan* plan, double* boost, double* refl) {
    BF_REQUIRE(nullptr, plan != nullptr, BF_ERR_ARG, "plan is NULL");
    BF_REQUIRE(plan->ctx, !refl || plan->has_reflect, BF_ERR_ARG, "refl requested but the plan has no reflect input");
    return guarded(plan->ctx, [&] { bf::transformer_plan_fetch(plan, boost, refl); });
}

int bf_transformer_plan_destroy(bf_transformer_plan* plan) {
    if (plan) bf::transformer_plan_destroy(plan);
    return BF_OK;
}

int bf_transformer_batch(bf_ctx* ctx, int32_t nconf, const bf_setup* setups, const bf_coords* coords, int32_t nsets,
                         const bf_modes* modesets, const double* const* emits, const int32_t* mode_set_of,
                         const double* freqs, int32_t nfreq, int32_t prop_kind, double diskR, int32_t precision,
                         const int64_t* boost_offset, double* boost) {
    TRY(check_kind(ctx, prop_kind));
    TRY(check_prec(ctx, precision));
    TRY(check_coords(ctx, coords));
    TRY(check_freqs(ctx, freqs, nfreq));
    BF_REQUIRE(ctx, nconf >= 0 && nsets >= 1, BF_ERR_ARG, "nconf must be >= 0 and nsets >= 1");
    BF_REQUIRE(ctx, setups && modesets && emits && mode_set_of && boost_offset && boost, BF_ERR_ARG,
               "setups, modesets, emits, mode_set_of, boost_offset and boost must not be NULL");
    for (int32_t m = 0; m < nsets; m++) {
        TRY(check_modes(ctx, &modesets[m]));
        BF_REQUIRE(ctx, emits[m] != nullptr, BF_ERR_ARG, "emits[m] is NULL");
    }
    for (int32_t c = 0; c < nconf; c++) {
        TRY(check_setup(ctx, &setups[c], 2, false));
        BF_REQUIRE(ctx, mode_set_of[c] >= 0 && mode_set_of[c] < nsets, BF_ERR_ARG, "mode_set_of[c] out of range");
        BF_REQUIRE(ctx, boost_offset[c] >= 0, BF_ERR_ARG, "boost_offset[c] is negative");
    }
    BF_REQUIRE(ctx, ctx != nullptr, BF_ERR_ARG, "ctx is NULL");
    return guarded(ctx, [&] {
        bf::transformer_batch_dev(ctx, nconf, setups, coords, nsets, modesets, emits, mode_set_of, freqs, nfreq,
                                  prop_kind, diskR, boost_offset, boost);
    });
}

// both plan constructors: one grid for every mode set (ncoords = 1) or one per mode set (ncoords = nsets)
static int batch_plan_create(bf_ctx* ctx, const bf_coords* coords, int32_t ncoords, int32_t nsets,
                             const bf_modes* modesets, const double* const* emits, const double* freqs, int32_t nfreq,
                             int32_t prop_kind, double diskR, int32_t precision, bf_transformer_batch_plan** out) {
    BF_REQUIRE(ctx, out != nullptr, BF_ERR_ARG, "out is NULL");
    *out = nullptr;
    TRY(check_kind(ctx, prop_kind));
    TRY(check_prec(ctx, precision));
    BF_REQUIRE(ctx, nsets >= 1, BF_ERR_ARG, "nsets must be >= 1");
    BF_REQUIRE(ctx, coords != nullptr, BF_ERR_ARG, "coords is NULL");
    for (int32_t k = 0; k < ncoords; k++) TRY(check_coords(ctx, &coords[k]));
    TRY(check_freqs(ctx, freqs, nfreq));
    BF_REQUIRE(ctx, modesets && emits, BF_ERR_ARG, "modesets and emits must not be NULL");
    for (int32_t m = 0; m < nsets; m++) {
        TRY(check_modes(ctx, &modesets[m]));
        BF_REQUIRE(ctx, emits[m] != nullptr, BF_ERR_ARG, "emits[m] is NULL");
    }
    BF_REQUIRE(ctx, ctx != nullptr, BF_ERR_ARG, "ctx is NULL");
    return guarded(ctx, [&] {
        *out = bf::transformer_batch_plan_create(ctx, coords, ncoords, nsets, modesets, emits, freqs, nfreq,
                                                 prop_kind, diskR);
    });
}

int bf_transformer_batch_plan_create(bf_ctx* ctx, const bf_coords* coords, int32_t nsets, const bf_modes* modesets,
                                     const double* const* emits, const double* freqs, int32_t nfreq,
                                     int32_t prop_kind, double diskR, int32_t precision,
                                     bf_transformer_batch_plan** out) {
    return batch_plan_create(ctx, coords, 1, nsets, modesets, emits, freqs, nfreq, prop_kind, diskR, precision, out);
}

int bf_transformer_batch_plan_create_grids(bf_ctx* ctx, int32_t nsets, const bf_coords* coords,
                                           const bf_modes* modesets, const double* const* emits, const double* freqs,
                                           int32_t nfreq, int32_t prop_kind, double diskR, int32_t precision,
                                           bf_transformer_batch_plan** out) {
    return batch_plan_create(ctx, coords, nsets, nsets, modesets, emits, freqs, nfreq, prop_kind, diskR, precision,
                             out);
}

int bf_transformer_batch_plan_run(bf_transformer_batch_plan* plan, int32_t nconf, const bf_setup* setups,
                                  const int32_t* mode_set_of, const int64_t* boost_offset) {
    BF_REQUIRE(nullptr, plan != nullptr, BF_ERR_ARG, "plan is NULL");
    bf_ctx* ctx = plan->ctx;
    const int32_t nsets = (int32_t)plan->md.size();
    BF_REQUIRE(ctx, nconf >= 0, BF_ERR_ARG, "nconf must be >= 0");
    BF_REQUIRE(ctx, nconf == 0 || (setups && mode_set_of && boost_offset), BF_ERR_ARG,
               "setups, mode_set_of and boost_offset must not be NULL");
    for (int32_t c = 0; c < nconf; c++) {
        TRY(check_setup(ctx, &setups[c], 2, false));
        BF_REQUIRE(ctx, mode_set_of[c] >= 0 && mode_set_of[c] < nsets, BF_ERR_ARG, "mode_set_of[c] out of range");
        BF_REQUIRE(ctx, boost_offset[c] >= 0, BF_ERR_ARG, "boost_offset[c] is negative");
    }
    return guarded(ctx, [&] { bf::transformer_batch_plan_run(plan, nconf, setups, mode_set_of, boost_offset); });
}

int bf_transformer_batch_plan_fetch(bf_transformer_batch_plan* plan, double* boost) {
    BF_REQUIRE(nullptr, plan != nullptr, BF_ERR_ARG, "plan is NULL");
    BF_REQUIRE(plan->ctx, boost != nullptr || plan->blocks.empty(), BF_ERR_ARG, "boost is NULL");
    return guarded(plan->ctx, [&] { bf::transformer_batch_plan_fetch(plan, boost); });
}

int bf_transformer_batch_plan_destroy(bf_transformer_batch_plan* plan) {
    if (plan) bf::transformer_batch_plan_destroy(plan);
    return BF_OK;
}

int bf_disk_system_1d(bf_ctx* ctx, const double* freqs, int32_t nfreq, int32_t ncand, int32_t num_disk,
                      const double* spacings, double disk_thickness, const double* eps_disk, const double* eps_gap,
                      int32_t mirror, const double* surfaceloss, double* gamma, double* axion) {
    TRY(check_freqs(ctx, freqs, nfreq));
    BF_REQUIRE(ctx, ncand >= 1 && num_disk >= 0, BF_ERR_ARG, "ncand >= 1 and num_disk >= 0 required");
    BF_REQUIRE(ctx, spacings && eps_disk && eps_gap && gamma && axion, BF_ERR_ARG,
               "spacings, eps_disk, eps_gap, gamma and axion must not be NULL");
    BF_REQUIRE(ctx, (int64_t)ncand * nfreq <= ((int64_t)1 << 31), BF_ERR_UNSUPPORTED, "more than 2^31 points");
    BF_REQUIRE(ctx, ctx != nullptr, BF_ERR_ARG, "ctx is NULL");
    return guarded(ctx, [&] {
        bf::disk_system_1d_dev(ctx, freqs, nfreq, ncand, num_disk, spacings, disk_thickness, eps_disk, eps_gap, mirror,
                               surfaceloss, gamma, axion);
    });
}

int bf_transform_surface_1d(bf_ctx* ctx, const double* freqs, int32_t n, const double* gamma_in, double* axion,
                            const double* eps_i, const double* eps_m, int32_t eps_m_real, double l,
                            double surfaceloss, int32_t use_ray, double* gamma_out) {
    TRY(check_freqs(ctx, freqs, n));
    BF_REQUIRE(ctx, gamma_in && eps_i && eps_m && gamma_out, BF_ERR_ARG,
               "gamma_in, eps_i, eps_m and gamma_out must not be NULL");
    BF_REQUIRE(ctx, use_ray || !axion, BF_ERR_ARG, "Need to use ray approx. (src/Analytical1D.py:44-45)");
    BF_REQUIRE(ctx, ctx != nullptr, BF_ERR_ARG, "ctx is NULL");
    return guarded(ctx, [&] {
        bf::transform_surface_1d_dev(ctx, freqs, n, gamma_in, axion, eps_i, eps_m, eps_m_real, l, surfaceloss, use_ray,
                                     gamma_out);
    });
}

int bf_propagation_matrices_grid(bf_ctx* ctx, const bf_setup* setup, const bf_coords* coords, const bf_modes* modes,
                                 const double* spacing_grid, int32_t nspacing, const double* tilt_x_grid,
                                 int32_t ntilt_x, const double* tilt_y_grid, int32_t ntilt_y, const double* freqs,
                                 int32_t nfreq, int32_t prop_kind, double diskR, double* out) {
    TRY(check_kind(ctx, prop_kind));
    TRY(check_coords(ctx, coords));
    TRY(check_setup(ctx, setup, 2, false));
    TRY(check_modes(ctx, modes));
    TRY(check_freqs(ctx, freqs, nfreq));
    BF_REQUIRE(ctx, setup->nreg % 2 == 0, BF_ERR_DIM,
               "the setup must have an even region count (mirror, gap, disk, ..., gap): "
               "examples/transformer_optim_utilities.jl:21-22");
    BF_REQUIRE(ctx, spacing_grid && tilt_x_grid && tilt_y_grid && out, BF_ERR_ARG,
               "spacing_grid, tilt_x_grid, tilt_y_grid and out must not be NULL");
    BF_REQUIRE(ctx, nspacing >= 1 && ntilt_x >= 1 && ntilt_y >= 1, BF_ERR_ARG, "grid lengths must be >= 1");
    BF_REQUIRE(ctx, ctx != nullptr, BF_ERR_ARG, "ctx is NULL");
    return guarded(ctx, [&] {
        bf::propagation_matrices_grid_dev(ctx, setup, coords, modes, spacing_grid, nspacing, tilt_x_grid, ntilt_x,
                                          tilt_y_grid, ntilt_y, freqs, nfreq, prop_kind, diskR, out);
    });
}

int bf_spacing_itp_create(bf_ctx* ctx, const double* grid, int32_t nreg, int32_t nfreq, int32_t nspacing, int32_t n,
                          double spacing_first, double spacing_step, bf_spacing_itp** out) {
    BF_REQUIRE(ctx, out != nullptr, BF_ERR_ARG, "out is NULL");
    *out = nullptr;
    BF_REQUIRE(ctx, grid != nullptr, BF_ERR_ARG, "grid is NULL");
    BF_REQUIRE(ctx, nreg >= 2 && nreg % 2 == 0, BF_ERR_DIM, "nreg must be even and >= 2");
    BF_REQUIRE(ctx, nfreq >= 1 && n >= 1 && n <= BF_MAX_MODES, BF_ERR_ARG, "nfreq >= 1 and 1 <= n <= 128 required");
    BF_REQUIRE(ctx, nspacing >= 2, BF_ERR_ARG, "the cubic spacing interpolation needs at least 2 grid points");
    BF_REQUIRE(ctx, std::isfinite(spacing_first) && std::isfinite(spacing_step) && spacing_step > 0.0, BF_ERR_ARG,
               "the spacing grid must be finite and increasing");
    BF_REQUIRE(ctx, ctx != nullptr, BF_ERR_ARG, "ctx is NULL");
    return guarded(ctx, [&] {
        *out = bf::spacing_itp_create(ctx, grid, nreg, nfreq, nspacing, n, spacing_first, spacing_step);
    });
}

int bf_spacing_itp_bounds(const bf_spacing_itp* itp, double* lo, double* hi) {
    BF_REQUIRE(nullptr, itp && lo && hi, BF_ERR_ARG, "itp, lo and hi must not be NULL");
    bf::spacing_itp_bounds(itp, lo, hi);
    return BF_OK;
}

int bf_spacing_itp_matrices(bf_spacing_itp* itp, int32_t ncand, const double* dist_shift, double* out) {
    BF_REQUIRE(nullptr, itp != nullptr, BF_ERR_ARG, "itp is NULL");
    bf_ctx* ctx = itp->ctx;
    BF_REQUIRE(ctx, ncand >= 0, BF_ERR_ARG, "ncand must be >= 0");
    BF_REQUIRE(ctx, ncand == 0 || (out && (dist_shift || itp->nreg == 2)), BF_ERR_ARG, "dist_shift or out is NULL");
    return guarded(ctx, [&] { bf::spacing_itp_matrices(itp, ncand, dist_shift, out); });
}

int bf_spacing_itp_transformer(bf_spacing_itp* itp, int32_t ncand, const double* dist_shift, const double* eps,
                               const double* emit, const double* reflect, double* boost, double* refl) {
    BF_REQUIRE(nullptr, itp != nullptr, BF_ERR_ARG, "itp is NULL");
    bf_ctx* ctx = itp->ctx;
    BF_REQUIRE(ctx, ncand >= 0, BF_ERR_ARG, "ncand must be >= 0");
    BF_REQUIRE(ctx, eps && emit && boost && (dist_shift || itp->nreg == 2 || ncand == 0), BF_ERR_ARG,
               "dist_shift, eps, emit and boost must not be NULL");
    BF_REQUIRE(ctx, !reflect || refl, BF_ERR_ARG, "refl output is NULL while reflect is given");
    return guarded(ctx, [&] { bf::spacing_itp_transformer(itp, ncand, dist_shift, eps, emit, reflect, boost, refl); });
}

int bf_spacing_itp_destroy(bf_spacing_itp* itp) {
    bf::spacing_itp_destroy(itp);
    return BF_OK;
}

int bf_transformer_trace_back(bf_ctx* ctx, const double* reflected, const double* input, const bf_setup* setup,
                              const bf_coords* coords, const bf_modes* modes, double f, int32_t prop_kind,
                              const double* prop_mats, double diskR, int32_t includes_axion, const double* emit,
                              double* fields) {
    TRY(check_kind(ctx, prop_kind));
    TRY(check_coords(ctx, coords));
    TRY(check_setup(ctx, setup, 2, false));
    TRY(check_modes(ctx, modes));
    TRY(check_freqs(ctx, &f, 1));
    BF_REQUIRE(ctx, reflected && input && fields, BF_ERR_ARG, "reflected, input and fields must not be NULL");
    BF_REQUIRE(ctx, !includes_axion || emit, BF_ERR_ARG, "emit is required when includes_axion is set");
    // the reference adds a 2 x n matrix to the 2n solution vector: only n == 1 conforms
    BF_REQUIRE(ctx, !includes_axion || modes->n == 1, BF_ERR_DIM,
               "includes_axion needs a single mode (2n vector + 2 x n matrix, src/boost_fractor_transformer.jl:411-413)");
    BF_REQUIRE(ctx, ctx != nullptr, BF_ERR_ARG, "ctx is NULL");
    return guarded(ctx, [&] {
        bf::trace_back_dev(ctx, reflected, input, setup, coords, modes, f, prop_kind, prop_mats, diskR, includes_axion,
                           emit, fields);
    });
}

int bf_transformer_chain_noalloc(bf_ctx* ctx, int32_t nreg, const double* eps, int32_t n, int32_t nfreq,
                                 const double* prop_mats, const double* emit, const double* reflect, double* boost,
                                 double* refl) {
    BF_REQUIRE(ctx, ctx != nullptr, BF_ERR_ARG, "ctx is NULL");
    BF_REQUIRE(ctx, eps && prop_mats && emit && boost, BF_ERR_ARG, "eps, prop_mats, emit and boost must not be NULL");
    BF_REQUIRE(ctx, !reflect == !refl, BF_ERR_ARG, "reflect and refl must be given together");
    BF_REQUIRE(ctx, nreg >= 2, BF_ERR_DIM,
               "the arrays in your SetupBoundaries objects don't fit together! (too few regions)");
    BF_REQUIRE(ctx, n >= 1 && nfreq >= 1, BF_ERR_ARG, "n and nfreq must be >= 1");
    BF_REQUIRE(ctx, n <= 48, BF_ERR_UNSUPPORTED, "the allocation-free chain holds n <= 48 modes");
    BF_REQUIRE(ctx, nreg <= BF_NOALLOC_MAX_REGIONS, BF_ERR_UNSUPPORTED,
               "the allocation-free chain holds at most BF_NOALLOC_MAX_REGIONS (512) regions");
    BF_REQUIRE(ctx, bf::noalloc_bytes(nreg, n, nfreq) <= (size_t)BF_NOALLOC_SCRATCH_BYTES, BF_ERR_UNSUPPORTED,
               "the call needs " + std::to_string(bf::noalloc_bytes(nreg, n, nfreq)) +
                   " bytes of working memory, above the allocation-free chain's BF_NOALLOC_SCRATCH_BYTES (32 MiB): "
                   "split the frequencies over several calls");
    return guarded(ctx, [&] {
        bf::transformer_chain_noalloc(ctx, nreg, eps, n, nfreq, reinterpret_cast<const zd*>(prop_mats),
                                      reinterpret_cast<const zd*>(emit), reinterpret_cast<const zd*>(reflect),
                                      reinterpret_cast<zd*>(boost), reinterpret_cast<zd*>(refl));
    });
}

}  // extern "C"
