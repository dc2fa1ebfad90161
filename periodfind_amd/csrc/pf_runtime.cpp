// Host runtime behind the C ABI (include/periodfind_amd.h).
//
// A pf_ctx owns one HIP stream on one device plus growable device buffers
// reused across calls (the reference allocates and frees everything per call,
// cuda/ce.cu:481-561).  A call uploads the CSR-packed curves once, prepares
// the derived per-point arrays the chosen statistic needs, launches one grid
// kernel over every (curve, 64-trial block), and copies the result back.
// Errors are returned as codes with a per-thread message instead of exit()
// (cuda/errchk.cuh:14-28).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/periodfind_amd.h"
#include "pf_kernels.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define PF_HIP(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(PF_E_HIP, "%s failed: %s", #call, hipGetErrorString(e_));        \
    } while (0)

#define PF_TRY(expr)             \
    do {                                 \
        int r_ = (expr);                 \
        if (r_ != PF_OK) return r_;      \
    } while (0)

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
};

// Grow-only device buffer.  Returns PF_OK or an error code.
int ensure(DevBuf& b, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (b.cap >= bytes) return PF_OK;
    if (b.p) {
        hipFree(b.p);
        b.p = nullptr;
        b.cap = 0;
    }
    size_t want = bytes + bytes / 4;
    if (std::getenv("PF_FAULT_ALLOC")) return fail(PF_E_NOMEM, "injected allocation failure");
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) {
        b.p = nullptr;
        return fail(PF_E_NOMEM, "hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
    }
    b.cap = want;
    return PF_OK;
}

template <typename T>
T* as(DevBuf& b) { return static_cast<T*>(b.p); }

}  // namespace

struct pf_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0.0f;
    size_t lds_max = 65536;
    // Inputs.
    DevBuf off, t, y, e;
    // Derived per-point arrays.
    DevBuf t64, tt, a0, a1, mrow;
    // Per-curve scalars.
    DevBuf cs0;
    // Trial grid (cached by content).
    DevBuf periods, pdts;
    std::vector<float> periods_h, pdts_h;
    // Output staging.
    DevBuf out, pidx, pval;
};

namespace {

struct Params {
    int algo;
    int n_phase, n_mag, phase_ext, mag_ext, n_bins;
    float qmin, qmax;
};

int check_params(const Params& p, size_t lds_max) {
    switch (p.algo) {
        case PF_ALGO_CE:
            if (p.n_phase < 1 || p.n_mag < 1 || p.phase_ext < 1 || p.mag_ext < 1)
                return fail(PF_E_ARG, "CE needs n_phase, n_mag, extents >= 1");
            if (pf::lds_ce(p.n_phase, p.n_mag) > lds_max)
                return fail(PF_E_LIMIT, "CE histogram %dx%d exceeds LDS (%zu B per wave)",
                            p.n_phase, p.n_mag, lds_max);
            break;
        case PF_ALGO_AOV:
            if (p.n_phase < 2 || p.phase_ext < 1)
                return fail(PF_E_ARG, "AOV needs n_phase >= 2 and phase_bin_extent >= 1");
            if (pf::lds_aov(p.n_phase) > lds_max)
                return fail(PF_E_LIMIT, "AOV n_phase=%d exceeds LDS", p.n_phase);
            break;
        case PF_ALGO_LS:
            break;
        case PF_ALGO_FPW:
            if (p.n_bins < 1) return fail(PF_E_ARG, "FPW needs n_bins >= 1");
            if (pf::lds_fpw(p.n_bins) > lds_max)
                return fail(PF_E_LIMIT, "FPW n_bins=%d exceeds LDS", p.n_bins);
            break;
        case PF_ALGO_BLS:
            if (p.n_bins < 2) return fail(PF_E_ARG, "BLS needs n_bins >= 2");
            if (!(p.qmin >= 0.0f) || !(p.qmax >= 0.0f))
                return fail(PF_E_ARG, "BLS needs qmin, qmax >= 0");
            if (pf::lds_bls(p.n_bins) > lds_max)
                return fail(PF_E_LIMIT, "BLS n_bins=%d exceeds LDS", p.n_bins);
            break;
        default:
            return fail(PF_E_ARG, "unknown algorithm %d", p.algo);
    }
    return PF_OK;
}

// rust/src/bls.rs:50-51: saturating f32 -> usize, then clamp.
void bls_duration_range(int nb, float qmin, float qmax, int* lo, int* hi) {
    auto sat = [](float x) -> uint64_t {
        if (!(x > 0.0f)) return 0;
        if (x >= 18446744073709551616.0f) return UINT64_MAX;
        return static_cast<uint64_t>(x);
    };
    uint64_t a = sat(std::floor(qmin * static_cast<float>(nb)));
    uint64_t b = sat(std::ceil(qmax * static_cast<float>(nb)));
    if (a < 1) a = 1;
    if (b > static_cast<uint64_t>(nb - 1)) b = static_cast<uint64_t>(nb - 1);
    *lo = static_cast<int>(a);
    *hi = static_cast<int>(b);
}

int upload_grid(pf_ctx* ctx, const float* periods, int64_t np, const float* pdts, int64_t nd,
                bool* has_pdt) {
    bool same = static_cast<int64_t>(ctx->periods_h.size()) == np &&
                static_cast<int64_t>(ctx->pdts_h.size()) == nd &&
                std::memcmp(ctx->periods_h.data(), periods, np * sizeof(float)) == 0 &&
                std::memcmp(ctx->pdts_h.data(), pdts, nd * sizeof(float)) == 0;
    *has_pdt = false;
    for (int64_t i = 0; i < nd; ++i)
        if (pdts[i] != 0.0f) *has_pdt = true;
    if (same) return PF_OK;
    PF_TRY(ensure(ctx->periods, np * sizeof(float)));
    PF_TRY(ensure(ctx->pdts, nd * sizeof(float)));
    PF_HIP(hipMemcpyAsync(ctx->periods.p, periods, np * sizeof(float), hipMemcpyHostToDevice,
                          ctx->stream));
    PF_HIP(hipMemcpyAsync(ctx->pdts.p, pdts, nd * sizeof(float), hipMemcpyHostToDevice,
                          ctx->stream));
    PF_HIP(hipStreamSynchronize(ctx->stream));
    ctx->periods_h.assign(periods, periods + np);
    ctx->pdts_h.assign(pdts, pdts + nd);
    return PF_OK;
}

// Curve set resident on the device for one call.
struct Curves {
    int64_t n_curves = 0;
    int64_t npts = 0;
    int64_t max_len = 0;
    const int64_t* off = nullptr;
    const float* t = nullptr;
    const float* y = nullptr;
    const float* e = nullptr;  // may be null (ones)
};

int validate_offsets(int64_t n_curves, const int64_t* off, int64_t* npts, int64_t* max_len) {
    if (n_curves < 0) return fail(PF_E_ARG, "n_curves < 0");
    if (!off) return fail(PF_E_ARG, "offsets is NULL");
    if (off[0] != 0) return fail(PF_E_ARG, "offsets[0] must be 0");
    int64_t mx = 0;
    for (int64_t c = 0; c < n_curves; ++c) {
        int64_t len = off[c + 1] - off[c];
        if (len < 0) return fail(PF_E_ARG, "offsets must be non-decreasing (curve %lld)",
                                 static_cast<long long>(c));
        if (len > mx) mx = len;
    }
    *npts = off[n_curves];
    *max_len = mx;
    return PF_OK;
}

int upload_curves(pf_ctx* ctx, int64_t n_curves, const int64_t* off, const float* t,
                  const float* y, const float* e, Curves* cv) {
    PF_TRY(validate_offsets(n_curves, off, &cv->npts, &cv->max_len));
    if (cv->npts > 0 && (!t || !y)) return fail(PF_E_ARG, "t and y must not be NULL");
    cv->n_curves = n_curves;
    const size_t nb = static_cast<size_t>(cv->npts) * sizeof(float);
    PF_TRY(ensure(ctx->off, (n_curves + 1) * sizeof(int64_t)));
    PF_TRY(ensure(ctx->t, nb));
    PF_TRY(ensure(ctx->y, nb));
    PF_HIP(hipMemcpyAsync(ctx->off.p, off, (n_curves + 1) * sizeof(int64_t),
                          hipMemcpyHostToDevice, ctx->stream));
    if (nb) {
        PF_HIP(hipMemcpyAsync(ctx->t.p, t, nb, hipMemcpyHostToDevice, ctx->stream));
        PF_HIP(hipMemcpyAsync(ctx->y.p, y, nb, hipMemcpyHostToDevice, ctx->stream));
    }
    cv->off = as<int64_t>(ctx->off);
    cv->t = as<float>(ctx->t);
    cv->y = as<float>(ctx->y);
    cv->e = nullptr;
    if (e) {
        PF_TRY(ensure(ctx->e, nb));
        if (nb) PF_HIP(hipMemcpyAsync(ctx->e.p, e, nb, hipMemcpyHostToDevice, ctx->stream));
        cv->e = as<float>(ctx->e);
    }
    return PF_OK;
}

// Prepare derived arrays and launch the grid kernel; out_d is [n_curves][T].
int run_grid(pf_ctx* ctx, const Params& p, const Curves& cv, int64_t np, int64_t nd,
             bool has_pdt, float* out_d) {
    PF_TRY(check_params(p, ctx->lds_max));
    const int64_t T = np * nd;
    if (cv.n_curves == 0 || T == 0) return PF_OK;
    hipStream_t s = ctx->stream;
    const int64_t n = cv.npts;

    PF_TRY(ensure(ctx->t64, n * sizeof(double)));
    if (has_pdt) PF_TRY(ensure(ctx->tt, n * sizeof(double)));
    PF_HIP(pf::launch_widen_times(cv.t, as<double>(ctx->t64), has_pdt ? as<double>(ctx->tt) : nullptr,
                                  n, s));

    pf::GridArgs g;
    g.off = cv.off;
    g.t64 = as<double>(ctx->t64);
    g.tt = has_pdt ? as<double>(ctx->tt) : as<double>(ctx->t64);
    g.periods = as<float>(ctx->periods);
    g.pdts = as<float>(ctx->pdts);
    g.nd = nd;
    g.T = T;
    g.tblocks = (T + pf::kWaveHost - 1) / pf::kWaveHost;
    g.n_blocks = static_cast<uint64_t>(cv.n_curves) * static_cast<uint64_t>(g.tblocks);
    g.out = out_d;
    g.has_pdt = has_pdt ? 1 : 0;
    if (g.n_blocks > 0xffffffffull)
        return fail(PF_E_LIMIT, "grid of %llu workgroups exceeds one launch",
                    static_cast<unsigned long long>(g.n_blocks));

    hipError_t le = hipSuccess;
    switch (p.algo) {
        case PF_ALGO_CE: {
            PF_TRY(ensure(ctx->mrow, n * sizeof(uint32_t)));
            PF_HIP(pf::launch_ce_mag_rows(cv.y, as<uint32_t>(ctx->mrow), n, p.n_mag, s));
            pf::CeArgs a{as<uint32_t>(ctx->mrow), p.n_phase, p.n_mag, p.phase_ext, p.mag_ext};
            PF_HIP(hipEventRecord(ctx->ev0, s));
            le = pf::launch_ce(g, a, pf::lds_ce(p.n_phase, p.n_mag), s);
            break;
        }
        case PF_ALGO_AOV: {
            PF_TRY(ensure(ctx->a0, n * sizeof(float)));
            PF_TRY(ensure(ctx->cs0, cv.n_curves * sizeof(float)));
            PF_HIP(pf::launch_square(cv.y, as<float>(ctx->a0), n, s));
            PF_HIP(pf::launch_curve_mean(cv.off, cv.y, as<float>(ctx->cs0), cv.n_curves, s));
            pf::AovArgs a{cv.y, as<float>(ctx->a0), as<float>(ctx->cs0), p.n_phase, p.phase_ext};
            PF_HIP(hipEventRecord(ctx->ev0, s));
            le = pf::launch_aov(g, a, pf::lds_aov(p.n_phase), s);
            break;
        }
        case PF_ALGO_LS: {
            pf::LsArgs a{cv.y, std::getenv("PF_LS_PRECISE") ? 0 : 1};
            PF_HIP(hipEventRecord(ctx->ev0, s));
            le = pf::launch_ls(g, a, s);
            break;
        }
        case PF_ALGO_FPW: {
            PF_TRY(ensure(ctx->a0, n * sizeof(float)));
            PF_TRY(ensure(ctx->a1, n * sizeof(float)));
            PF_HIP(pf::launch_inv_var(cv.e, as<float>(ctx->a0), n, s));
            PF_HIP(pf::launch_mul(as<float>(ctx->a0), cv.y, as<float>(ctx->a1), n, s));
            pf::FpwArgs a{as<float>(ctx->a0), as<float>(ctx->a1), p.n_bins};
            PF_HIP(hipEventRecord(ctx->ev0, s));
            le = pf::launch_fpw(g, a, pf::lds_fpw(p.n_bins), s);
            break;
        }
        case PF_ALGO_BLS: {
            PF_TRY(ensure(ctx->a0, n * sizeof(float)));
            PF_TRY(ensure(ctx->a1, n * sizeof(float)));
            PF_TRY(ensure(ctx->cs0, cv.n_curves * sizeof(float)));
            PF_HIP(pf::launch_inv_var(cv.e, as<float>(ctx->a0), n, s));
            PF_HIP(pf::launch_bls_prepare(cv.off, cv.y, as<float>(ctx->a0), as<float>(ctx->cs0),
                                          as<float>(ctx->a1), cv.n_curves, cv.max_len, s));
            int lo, hi;
            bls_duration_range(p.n_bins, p.qmin, p.qmax, &lo, &hi);
            pf::BlsArgs a{as<float>(ctx->a0), as<float>(ctx->a1), p.n_bins, lo, hi};
            PF_HIP(hipEventRecord(ctx->ev0, s));
            le = pf::launch_bls(g, a, pf::lds_bls(p.n_bins), s);
            break;
        }
    }
    if (le != hipSuccess) return fail(PF_E_HIP, "grid kernel launch failed: %s", hipGetErrorString(le));
    PF_HIP(hipEventRecord(ctx->ev1, s));
    return PF_OK;
}

int finish_timing(pf_ctx* ctx) {
    PF_HIP(hipEventSynchronize(ctx->ev1));
    float ms = 0.0f;
    PF_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_ms = ms;
    return PF_OK;
}

Params to_params(const pf_params* p) {
    Params q;
    q.algo = p->algo;
    q.n_phase = p->n_phase;
    q.n_mag = p->n_mag;
    q.phase_ext = p->phase_ext;
    q.mag_ext = p->mag_ext;
    q.n_bins = p->n_bins;
    q.qmin = p->qmin;
    q.qmax = p->qmax;
    return q;
}

int check_grid(const float* periods, int64_t np, const float* pdts, int64_t nd) {
    if (np < 0 || nd < 0) return fail(PF_E_ARG, "negative grid size");
    if ((np > 0 && !periods) || (nd > 0 && !pdts)) return fail(PF_E_ARG, "grid pointer is NULL");
    return PF_OK;
}

// Host in -> host out grid evaluation shared by the per-algorithm entry points.
int grid_host(pf_ctx* ctx, const Params& p, int64_t n_curves, const int64_t* off, const float* t,
              const float* y, const float* e, const float* periods, int64_t np,
              const float* pdts, int64_t nd, float* out) {
    if (!ctx) return fail(PF_E_ARG, "ctx is NULL");
    PF_TRY(check_params(p, ctx->lds_max));
    PF_TRY(check_grid(periods, np, pdts, nd));
    if (n_curves > 0 && !out) return fail(PF_E_ARG, "out is NULL");
    PF_HIP(hipSetDevice(ctx->device));
    Curves cv;
    PF_TRY(upload_curves(ctx, n_curves, off, t, y, e, &cv));
    const int64_t T = np * nd;
    if (n_curves == 0) return PF_OK;
    if (T == 0) return PF_OK;
    bool has_pdt = false;
    PF_TRY(upload_grid(ctx, periods, np, pdts, nd, &has_pdt));
    const size_t ob = static_cast<size_t>(n_curves) * T * sizeof(float);
    PF_TRY(ensure(ctx->out, ob));
    PF_TRY(run_grid(ctx, p, cv, np, nd, has_pdt, as<float>(ctx->out)));
    PF_HIP(hipMemcpyAsync(out, ctx->out.p, ob, hipMemcpyDeviceToHost, ctx->stream));
    PF_HIP(hipStreamSynchronize(ctx->stream));
    return finish_timing(ctx);
}

}  // namespace

extern "C" {

const char* pf_last_error(void) { return g_err.c_str(); }

int pf_device_count(int* count) {
    if (!count) return fail(PF_E_ARG, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return PF_OK;
}

int pf_ctx_create(pf_ctx** out, int device) {
    if (!out) return fail(PF_E_ARG, "ctx out-pointer is NULL");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return fail(PF_E_NODEVICE, "no HIP device available");
    if (device < 0 || device >= n) return fail(PF_E_ARG, "device %d out of range [0,%d)", device, n);
    PF_HIP(hipSetDevice(device));
    pf_ctx* ctx = new pf_ctx();
    ctx->device = device;
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) ==
            hipSuccess && lds > 0)
        ctx->lds_max = static_cast<size_t>(lds);
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) {
        pf_ctx_destroy(ctx);
        return fail(PF_E_HIP, "stream/event creation failed");
    }
    *out = ctx;
    return PF_OK;
}

void pf_ctx_destroy(pf_ctx* ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    DevBuf* bufs[] = {&ctx->off, &ctx->t, &ctx->y, &ctx->e, &ctx->t64, &ctx->tt, &ctx->a0,
                      &ctx->a1, &ctx->mrow, &ctx->cs0, &ctx->periods, &ctx->pdts, &ctx->out,
                      &ctx->pidx, &ctx->pval};
    for (DevBuf* b : bufs)
        if (b->p) hipFree(b->p);
    if (ctx->ev0) hipEventDestroy(ctx->ev0);
    if (ctx->ev1) hipEventDestroy(ctx->ev1);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
}

int pf_ce_grid(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets, const float* t,
               const float* y, const float* periods, int64_t n_periods,
               const float* period_dts, int64_t n_pdts, int n_phase, int n_mag, int phase_ext,
               int mag_ext, float* out) {
    Params p{PF_ALGO_CE, n_phase, n_mag, phase_ext, mag_ext, 0, 0.0f, 0.0f};
    return grid_host(ctx, p, n_curves, offsets, t, y, nullptr, periods, n_periods, period_dts,
                     n_pdts, out);
}

int pf_aov_grid(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets, const float* t,
                const float* y, const float* periods, int64_t n_periods,
                const float* period_dts, int64_t n_pdts, int n_phase, int phase_ext,
                float* out) {
    Params p{PF_ALGO_AOV, n_phase, 0, phase_ext, 1, 0, 0.0f, 0.0f};
    return grid_host(ctx, p, n_curves, offsets, t, y, nullptr, periods, n_periods, period_dts,
                     n_pdts, out);
}

int pf_ls_grid(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets, const float* t,
               const float* y, const float* periods, int64_t n_periods,
               const float* period_dts, int64_t n_pdts, float* out) {
    Params p{PF_ALGO_LS, 0, 0, 1, 1, 0, 0.0f, 0.0f};
    return grid_host(ctx, p, n_curves, offsets, t, y, nullptr, periods, n_periods, period_dts,
                     n_pdts, out);
}

int pf_fpw_grid(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets, const float* t,
                const float* y, const float* err, const float* periods, int64_t n_periods,
                const float* period_dts, int64_t n_pdts, int n_bins, float* out) {
    Params p{PF_ALGO_FPW, 0, 0, 1, 1, n_bins, 0.0f, 0.0f};
    return grid_host(ctx, p, n_curves, offsets, t, y, err, periods, n_periods, period_dts,
                     n_pdts, out);
}

int pf_bls_grid(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets, const float* t,
                const float* y, const float* err, const float* periods, int64_t n_periods,
                const float* period_dts, int64_t n_pdts, int n_bins, float qmin, float qmax,
                float* out) {
    Params p{PF_ALGO_BLS, 0, 0, 1, 1, n_bins, qmin, qmax};
    return grid_host(ctx, p, n_curves, offsets, t, y, err, periods, n_periods, period_dts,
                     n_pdts, out);
}

int pf_peaks(pf_ctx* ctx, const pf_params* params, int64_t n_curves, const int64_t* offsets,
             const float* t, const float* y, const float* err, const float* periods,
             int64_t n_periods, const float* period_dts, int64_t n_pdts, int64_t n_peaks,
             int64_t min_distance, int64_t* idx, float* val) {
    if (!ctx || !params) return fail(PF_E_ARG, "ctx/params is NULL");
    if (n_peaks < 0 || n_peaks > 64) return fail(PF_E_LIMIT, "n_peaks must be in [0, 64]");
    if (min_distance < 0) return fail(PF_E_ARG, "min_distance < 0");
    Params p = to_params(params);
    PF_TRY(check_params(p, ctx->lds_max));
    PF_TRY(check_grid(periods, n_periods, period_dts, n_pdts));
    if (n_curves > 0 && n_peaks > 0 && (!idx || !val)) return fail(PF_E_ARG, "idx/val is NULL");
    PF_HIP(hipSetDevice(ctx->device));
    Curves cv;
    PF_TRY(upload_curves(ctx, n_curves, offsets, t, y, err, &cv));
    if (n_curves == 0 || n_peaks == 0) return PF_OK;
    const int64_t T = n_periods * n_pdts;
    const size_t pk = static_cast<size_t>(n_curves) * n_peaks;
    if (T == 0) {
        for (size_t i = 0; i < pk; ++i) { idx[i] = -1; val[i] = 0.0f; }
        return PF_OK;
    }
    bool has_pdt = false;
    PF_TRY(upload_grid(ctx, periods, n_periods, period_dts, n_pdts, &has_pdt));
    PF_TRY(ensure(ctx->out, static_cast<size_t>(n_curves) * T * sizeof(float)));
    PF_TRY(ensure(ctx->pidx, pk * sizeof(int64_t)));
    PF_TRY(ensure(ctx->pval, pk * sizeof(float)));
    PF_TRY(run_grid(ctx, p, cv, n_periods, n_pdts, has_pdt, as<float>(ctx->out)));
    const int use_max = p.algo == PF_ALGO_CE ? 0 : 1;
    PF_HIP(pf::launch_peaks(as<float>(ctx->out), n_curves, T, n_peaks, min_distance, use_max,
                            as<int64_t>(ctx->pidx), as<float>(ctx->pval), ctx->stream));
    PF_HIP(hipMemcpyAsync(idx, ctx->pidx.p, pk * sizeof(int64_t), hipMemcpyDeviceToHost,
                          ctx->stream));
    PF_HIP(hipMemcpyAsync(val, ctx->pval.p, pk * sizeof(float), hipMemcpyDeviceToHost,
                          ctx->stream));
    PF_HIP(hipStreamSynchronize(ctx->stream));
    return finish_timing(ctx);
}

int pf_find_top_peaks(pf_ctx* ctx, const float* data, int64_t n_curves, int64_t T,
                      int64_t n_peaks, int64_t min_distance, int use_max, int64_t* idx,
                      float* val) {
    if (!ctx) return fail(PF_E_ARG, "ctx is NULL");
    if (n_curves < 0 || T < 0) return fail(PF_E_ARG, "negative size");
    if (n_peaks < 0 || n_peaks > 64) return fail(PF_E_LIMIT, "n_peaks must be in [0, 64]");
    if (min_distance < 0) return fail(PF_E_ARG, "min_distance < 0");
    if (n_curves == 0 || n_peaks == 0) return PF_OK;
    if (!idx || !val || (T > 0 && !data)) return fail(PF_E_ARG, "NULL pointer");
    PF_HIP(hipSetDevice(ctx->device));
    const size_t pk = static_cast<size_t>(n_curves) * n_peaks;
    const size_t db = static_cast<size_t>(n_curves) * T * sizeof(float);
    PF_TRY(ensure(ctx->out, db));
    PF_TRY(ensure(ctx->pidx, pk * sizeof(int64_t)));
    PF_TRY(ensure(ctx->pval, pk * sizeof(float)));
    if (db) PF_HIP(hipMemcpyAsync(ctx->out.p, data, db, hipMemcpyHostToDevice, ctx->stream));
    if (T == 0) {
        for (size_t i = 0; i < pk; ++i) { idx[i] = -1; val[i] = 0.0f; }
        return PF_OK;
    }
    PF_HIP(pf::launch_peaks(as<float>(ctx->out), n_curves, T, n_peaks, min_distance,
                            use_max ? 1 : 0, as<int64_t>(ctx->pidx), as<float>(ctx->pval),
                            ctx->stream));
    PF_HIP(hipMemcpyAsync(idx, ctx->pidx.p, pk * sizeof(int64_t), hipMemcpyDeviceToHost,
                          ctx->stream));
    PF_HIP(hipMemcpyAsync(val, ctx->pval.p, pk * sizeof(float), hipMemcpyDeviceToHost,
                          ctx->stream));
    PF_HIP(hipStreamSynchronize(ctx->stream));
    return PF_OK;
}

int pf_grid_device(pf_ctx* ctx, const pf_params* params, int64_t n_curves,
                   const int64_t* offsets_d, const float* t_d, const float* y_d,
                   const float* err_d, int64_t max_len, const float* periods_d,
                   int64_t n_periods, const float* period_dts_d, int64_t n_pdts, float* out_d,
                   int sync) {
    (void)max_len;
    (void)periods_d;
    (void)period_dts_d;
    (void)ctx; (void)params; (void)n_curves; (void)offsets_d; (void)t_d; (void)y_d;
    (void)err_d; (void)n_periods; (void)n_pdts; (void)out_d; (void)sync;
    return fail(PF_E_ARG, "pf_grid_device: not implemented");
}

int pf_last_kernel_ms(pf_ctx* ctx, float* ms) {
    if (!ctx || !ms) return fail(PF_E_ARG, "NULL argument");
    *ms = ctx->last_ms;
    return PF_OK;
}

}  // extern "C"
