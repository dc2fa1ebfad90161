// Host runtime behind the C ABI (include/periodfind_amd.h).
//
// A pf_ctx owns one HIP stream on one device plus growable device buffers
// reused across calls (the reference allocates and frees everything per call,
// cuda/ce.cu:481-561).  A call uploads the CSR-packed curves once, prepares
// the derived per-point arrays the chosen statistic needs, launches one grid
// kernel over every (curve, 64-trial block), and copies the result back.
// Errors are returned as codes with a per-thread message instead of exit()
// (cuda/errchk.cuh:14-28).
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
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
    bool fixed = false;  // a region of memory the buffer does not own: never regrown
};

// Grow-only device buffer.  Returns PF_OK or an error code.
int ensure(DevBuf& b, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (b.cap >= bytes) return PF_OK;
    if (b.fixed)
        return fail(PF_E_LIMIT, "a fixed %zu-byte region cannot hold %zu bytes", b.cap, bytes);
    if (b.p) {
        (void)hipFree(b.p);
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

// A curve set on the device: owned CSR storage plus the derived per-point
// arrays each statistic needs, built on first use and kept until the inputs
// change.  The drop-in calls reuse one set owned by the context; resident
// sets (pf_curves) own theirs.
struct CurveSet {
    int64_t n_curves = 0;
    int64_t npts = 0;
    int64_t max_len = 0;
    bool has_err = false;
    DevBuf off, t, y, e;
    // Prepared magnitudes (PF_MAG_CENTER / PF_MAG_NORMALIZE of y) and the
    // mode they hold (-1: none); derived arrays below were built from the
    // magnitudes of mode `derived_mode`.
    DevBuf yp;
    int yp_mode = -1;
    int derived_mode = PF_MAG_AS_IS;
    // Derived, with validity flags.  tmax (per curve, the largest |t|) is
    // built with t64.
    DevBuf t64, tt, tmax;
    bool t64_ok = false, tt_ok = false;
    DevBuf ce_mrow, ce_mpk;  // mpk: built with mrow for an even n_mag
    int ce_nmag = -1;
    DevBuf aov_yy, aov_avg;
    bool aov_ok = false;
    DevBuf w;  // 1/(e*e) (FPW and BLS)
    bool w_ok = false;
    DevBuf fpw_wy;
    bool fpw_ok = false;
    DevBuf bls_yw, bls_mean;
    bool bls_ok = false;
    // Equal weights: running serial sums of each curve's first weight, and a
    // device flag, nonzero unless every curve's weights are bitwise equal
    // (then FPW and BLS count points per bin and weigh a bin of c points by
    // wsum[c - 1]).  The flag's value on the host once read back (BLS picks
    // its kernel on the host).  Built from w alone, so kept across magnitude
    // modes.
    DevBuf wsum, w_flag;
    bool wsum_ok = false, w_equal_known = false, w_equal = false;
    DevBuf mf_wyy;  // (w*y)*y (MF); w and w*y are shared with FPW
    bool mf_ok = false;
    DevBuf vn_lut, vn_lut_off;  // VN emission tables, for extents product vn_ext
    int vn_ext = -1;
    std::vector<int64_t> vn_lut_off_h;  // host copy (source of the async upload)
    DevBuf mhf_pt, mhf_sums;  // MHF point records and per-curve sums
    bool mhf_ok = false;
    DevBuf mhf_part;          // MHF partial sums of split launches
    // Last periodogram.
    DevBuf out;
    int64_t out_T = 0;
    // Ragged (persistent) launches: curves longest first, and the work counter.
    std::vector<int64_t> off_h;  // host copy of the offsets
    int64_t rep_len = 0;         // work-weighted mean length sum(len^2)/sum(len)
    std::vector<int64_t> order_h;  // curves longest first (host copy of `order`)
    DevBuf order, counter;
    bool order_ok = false;
    // Captured launch sequence of the ragged peaks path, and what it was built for.
    hipGraphExec_t graph = nullptr;
    std::vector<uint64_t> graph_key;

    void drop_graph() {
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
        graph_key.clear();
    }
    void invalidate() {
        t64_ok = tt_ok = aov_ok = w_ok = fpw_ok = bls_ok = mf_ok = mhf_ok = order_ok = false;
        wsum_ok = w_equal_known = w_equal = false;
        vn_ext = -1;
        ce_nmag = -1;
        out_T = 0;
        yp_mode = -1;
        derived_mode = PF_MAG_AS_IS;
        drop_graph();
    }
    // Drop what was built from the magnitudes (a new preparation mode).
    void invalidate_mags() {
        aov_ok = fpw_ok = bls_ok = mf_ok = mhf_ok = false;
        ce_nmag = -1;
        drop_graph();
    }
    std::vector<DevBuf*> buffers() {
        return {&off, &t, &y, &e, &yp, &t64, &tt, &tmax, &ce_mrow, &ce_mpk, &aov_yy, &aov_avg, &w,
                &fpw_wy, &bls_yw, &bls_mean, &wsum, &w_flag, &mf_wyy, &vn_lut,
                &vn_lut_off, &mhf_pt, &mhf_sums, &mhf_part, &out, &order, &counter};
    }
    // A set over memory it does not own: no buffer may grow (a grow fails
    // instead of allocating).
    void fix_all() {
        for (DevBuf* b : buffers()) b->fixed = true;
    }
    void release() {
        drop_graph();
        for (DevBuf* b : buffers())
            if (b->p && !b->fixed) {
                (void)hipFree(b->p);
                b->p = nullptr;
                b->cap = 0;
            }
        invalidate();
    }
};

}  // namespace

struct pf_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timed = false;
    float last_ms = 0.0f;
    size_t lds_max = 65536;
    int n_cu = 256;  // compute units of the device
    int rmw_group = 2;  // LDS read-modify-write group size of the f32-binned kernels
    int ce_group = 0;   // CE: 0 = LDS u32 atomics, else read-modify-write group size
    // Per-(statistic, length bucket) launch choices; empty = the defaults above.
    std::vector<pf_tune_entry> tune;
    std::string tune_source;
    std::string arch;  // e.g. "gfx950"
    CurveSet scratch;  // curve set of the drop-in calls
    // Trial grid (cached by content).
    DevBuf periods, pdts;
    std::vector<float> periods_h, pdts_h;
    bool has_pdt = false;
    // Outputs of one call, back to back (see OutPlan), and their pinned host
    // staging buffer.
    DevBuf outbuf;
    void* host_stage = nullptr;
    size_t host_cap = 0;
    // Input staging of pf_curves_load: two pinned chunks, each reused once
    // the copy issued from it has completed (its event).
    void* in_stage[2] = {nullptr, nullptr};
    hipEvent_t in_ev[2] = {nullptr, nullptr};
    bool in_busy[2] = {false, false};
    // Side streams of pf_curves_stats_many (default priority), created on
    // first use, and a start/end event pair per job.
    static constexpr int kSideStreams = 3;
    hipStream_t side[kSideStreams] = {nullptr, nullptr, nullptr};
    hipEvent_t join = nullptr;
    std::vector<hipEvent_t> job_ev;
    // LS refinement: frequency grid (f64) and per-peak brackets.
    DevBuf ls_freq, ls_lo, ls_hi;
    // Segment selections of a split peak scan (see pf::launch_peaks).
    DevBuf peaks_ws;
    // High-cadence filter: kept indices, counts, a copy of the set's planes
    // and the new offsets.
    DevBuf hc_keep, hc_cnt, hc_src, hc_off;
    // Matched Filter templates, one table per bin count (rows of
    // mf_template_stride(nb)); a call with several MF jobs of different bin
    // counts reads each job's own table.
    struct MfTable {
        int nb;
        DevBuf buf;
        std::vector<float> h;  // host copy (source of the async upload)
    };
    std::vector<MfTable> mf_tables;
    // The device array of pf_peaks_noalloc on this device (resolved at
    // creation) and the status of that lookup.
    void* noalloc = nullptr;
    hipError_t noalloc_rc = hipErrorNotInitialized;
};

struct pf_curves {
    pf_ctx* ctx = nullptr;
    CurveSet cs;
};

namespace {

struct Params {
    int algo;
    int n_phase, n_mag, phase_ext, mag_ext, n_bins;
    float qmin, qmax;
    int mag_prep;  // PF_MAG_*
    int bandwidth, margin;  // VN
    int max_harmonics;      // MHF
};

// --- Launch tuning -----------------------------------------------------------
// Results do not depend on these choices (every variant accumulates in the
// same order); only speed does.

struct Choice {
    int group;  // CE: 0 = LDS u32 atomics, 1/2/4 = read-modify-write group; others 1/2/4
    int claim;  // work items per atomic claim of a persistent launch
};

Choice tune_choice(const pf_ctx* ctx, int algo, int64_t len) {
    Choice c{algo == PF_ALGO_CE ? ctx->ce_group : ctx->rmw_group, 16};
    int64_t best = -1;
    for (const pf_tune_entry& e : ctx->tune)
        if (e.algo == algo && e.min_len <= len && e.min_len > best) {
            best = e.min_len;
            c = Choice{e.group, e.claim};
        }
    return c;
}

// Ascending bucket lower bounds of `algo` (always starting at 0).
std::vector<int64_t> tune_edges(const pf_ctx* ctx, int algo) {
    std::vector<int64_t> ed{0};
    for (const pf_tune_entry& e : ctx->tune)
        if (e.algo == algo && e.min_len > 0) ed.push_back(e.min_len);
    std::sort(ed.begin(), ed.end());
    ed.erase(std::unique(ed.begin(), ed.end()), ed.end());
    return ed;
}

constexpr int kMaxTuneBuckets = 32;
const char* const kAlgoNames[] = {"ce", "aov", "ls", "fpw", "bls", "mf", "vn", "mhf"};
constexpr int kNumAlgos = 8;

int validate_tune(const pf_tune_entry* t, int64_t n) {
    if (n < 0 || (n > 0 && !t)) return fail(PF_E_ARG, "tuning table pointer/size invalid");
    int per_algo[kNumAlgos] = {};
    for (int64_t i = 0; i < n; ++i) {
        const pf_tune_entry& e = t[i];
        if (e.algo < PF_ALGO_CE || e.algo >= kNumAlgos)
            return fail(PF_E_ARG, "tuning entry %lld: unknown algorithm %d",
                        static_cast<long long>(i), e.algo);
        if (e.min_len < 0)
            return fail(PF_E_ARG, "tuning entry %lld: min_len < 0", static_cast<long long>(i));
        if (e.claim < 1 || e.claim > 64)
            return fail(PF_E_ARG, "tuning entry %lld: claim must be in [1, 64]",
                        static_cast<long long>(i));
        const bool g124 = e.group == 1 || e.group == 2 || e.group == 4;
        if (!(g124 || (e.algo == PF_ALGO_CE && e.group == 0)))
            return fail(PF_E_ARG, "tuning entry %lld: group %d not supported for %s",
                        static_cast<long long>(i), e.group, kAlgoNames[e.algo]);
        // Plus the implicit bucket from length 0: at most kMaxTuneBuckets in all.
        if (++per_algo[e.algo] > kMaxTuneBuckets - 1)
            return fail(PF_E_LIMIT, "more than %d tuning entries for %s", kMaxTuneBuckets - 1,
                        kAlgoNames[e.algo]);
    }
    return PF_OK;
}

// Text table: one entry per line, "<algo> <min_len> <group> <claim>",
// algo in {ce, aov, ls, fpw, bls, mf, vn, mhf}; '#' starts a comment.
int parse_tune_file(const char* path, std::vector<pf_tune_entry>* out) {
    std::ifstream f(path);
    if (!f) return fail(PF_E_ARG, "cannot open tuning table %s", path);
    std::string line;
    int lineno = 0;
    out->clear();
    while (std::getline(f, line)) {
        ++lineno;
        const size_t h = line.find('#');
        if (h != std::string::npos) line.resize(h);
        std::istringstream ss(line);
        std::string name;
        if (!(ss >> name)) continue;
        pf_tune_entry e{};
        e.algo = -1;
        for (int a = 0; a < kNumAlgos; ++a)
            if (name == kAlgoNames[a]) e.algo = a;
        long long ml = 0;
        if (e.algo < 0 || !(ss >> ml >> e.group >> e.claim))
            return fail(PF_E_ARG, "%s:%d: expected '<algo> <min_len> <group> <claim>'", path,
                        lineno);
        e.min_len = ml;
        out->push_back(e);
    }
    return validate_tune(out->data(), static_cast<int64_t>(out->size()));
}

// Base architecture name of a device ("gfx950" of "gfx950:sramecc+:xnack-").
std::string device_arch(int device) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return "";
    std::string arch = prop.gcnArchName;
    const size_t colon = arch.find(':');
    if (colon != std::string::npos) arch.resize(colon);
    return arch;
}

// The table shipped next to this library for an architecture
// (tuning_<arch>.txt), or "" when there is none.
std::string default_tune_path(const std::string& arch) {
    if (arch.empty()) return "";
    Dl_info info;
    if (!dladdr(reinterpret_cast<void*>(&default_tune_path), &info) || !info.dli_fname) return "";
    std::string dir = info.dli_fname;
    const size_t slash = dir.rfind('/');
    dir = slash == std::string::npos ? std::string(".") : dir.substr(0, slash);
    std::string path = dir + "/tuning_" + arch + ".txt";
    std::ifstream f(path);
    return f ? path : "";
}

int check_params(const Params& p, size_t lds_max) {
    if (p.mag_prep != PF_MAG_AS_IS && p.mag_prep != PF_MAG_CENTER && p.mag_prep != PF_MAG_NORMALIZE)
        return fail(PF_E_ARG, "mag_prep must be PF_MAG_AS_IS, PF_MAG_CENTER or PF_MAG_NORMALIZE");
    switch (p.algo) {
        case PF_ALGO_CE:
            if (p.n_phase < 1 || p.n_mag < 1 || p.phase_ext < 1 || p.mag_ext < 1)
                return fail(PF_E_ARG, "CE needs n_phase, n_mag, extents >= 1");
            if (pf::lds_ce(p.n_phase, p.n_mag) + pf::kCeStaticLds > lds_max)
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
        case PF_ALGO_VN:
            if (p.n_phase < 1 || p.n_mag < 1 || p.phase_ext < 1 || p.mag_ext < 1)
                return fail(PF_E_ARG, "VN needs n_phase, n_mag, extents >= 1");
            if (p.bandwidth < 0 || p.margin < 0)
                return fail(PF_E_ARG, "VN needs bandwidth, margin >= 0");
            if (p.n_mag > pf::kVnMaxMag)
                return fail(PF_E_LIMIT, "VN n_mag=%d exceeds %d", p.n_mag, pf::kVnMaxMag);
            if (pf::lds_ce(p.n_phase, p.n_mag) > lds_max)
                return fail(PF_E_LIMIT, "VN histogram %dx%d exceeds LDS (%zu B per wave)",
                            p.n_phase, p.n_mag, lds_max);
            break;
        case PF_ALGO_MHF:
            if (p.max_harmonics < 0) return fail(PF_E_ARG, "MHF needs max_harmonics >= 0");
            break;
        case PF_ALGO_MF:
            if (p.n_bins < 1) return fail(PF_E_ARG, "MF needs num_bins >= 1");
            if (pf::lds_mf(p.n_bins) > lds_max)
                return fail(PF_E_LIMIT, "MF num_bins=%d exceeds LDS", p.n_bins);
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
    // Any a above b searches nothing; nb keeps it above b and within int.
    if (a > static_cast<uint64_t>(nb)) a = static_cast<uint64_t>(nb);
    *lo = static_cast<int>(a);
    *hi = static_cast<int>(b);
}

int check_grid(const float* periods, int64_t np, const float* pdts, int64_t nd) {
    if (np < 0 || nd < 0) return fail(PF_E_ARG, "negative grid size");
    if ((np > 0 && !periods) || (nd > 0 && !pdts)) return fail(PF_E_ARG, "grid pointer is NULL");
    return PF_OK;
}

int upload_grid(pf_ctx* ctx, const float* periods, int64_t np, const float* pdts, int64_t nd) {
    bool same = static_cast<int64_t>(ctx->periods_h.size()) == np &&
                static_cast<int64_t>(ctx->pdts_h.size()) == nd &&
                std::memcmp(ctx->periods_h.data(), periods, np * sizeof(float)) == 0 &&
                std::memcmp(ctx->pdts_h.data(), pdts, nd * sizeof(float)) == 0;
    if (same) return PF_OK;
    // Forget the cached grid first: a failed grow below frees the old buffers.
    ctx->periods_h.clear();
    ctx->pdts_h.clear();
    bool has_pdt = false;
    for (int64_t i = 0; i < nd; ++i)
        if (pdts[i] != 0.0f) has_pdt = true;
    PF_TRY(ensure(ctx->periods, np * sizeof(float)));
    PF_TRY(ensure(ctx->pdts, nd * sizeof(float)));
    PF_HIP(hipMemcpyAsync(ctx->periods.p, periods, np * sizeof(float), hipMemcpyHostToDevice,
                          ctx->stream));
    PF_HIP(hipMemcpyAsync(ctx->pdts.p, pdts, nd * sizeof(float), hipMemcpyHostToDevice,
                          ctx->stream));
    PF_HIP(hipStreamSynchronize(ctx->stream));
    ctx->periods_h.assign(periods, periods + np);
    ctx->pdts_h.assign(pdts, pdts + nd);
    ctx->has_pdt = has_pdt;
    return PF_OK;
}

int validate_offsets(int64_t n_curves, const int64_t* off, int64_t* npts, int64_t* max_len) {
    if (n_curves < 0) return fail(PF_E_ARG, "n_curves < 0");
    if (!off) return fail(PF_E_ARG, "offsets is NULL");
    if (off[0] != 0) return fail(PF_E_ARG, "offsets[0] must be 0");
    int64_t mx = 0;
    for (int64_t c = 0; c < n_curves; ++c) {
        int64_t len = off[c + 1] - off[c];
        if (len < 0)
            return fail(PF_E_ARG, "offsets must be non-decreasing (curve %lld)",
                        static_cast<long long>(c));
        if (len > mx) mx = len;
    }
    *npts = off[n_curves];
    *max_len = mx;
    return PF_OK;
}

// Host per-curve arrays -> one device CSR array, through the context's two
// pinned staging chunks: a chunk is filled on the host and its copy issued
// on the stream, then the other chunk is filled while that copy runs; a
// chunk is refilled only once its previous copy has completed (its event).
constexpr size_t kInChunk = size_t(4) << 20;

int stage_init(pf_ctx* ctx) {
    for (int i = 0; i < 2; ++i) {
        if (!ctx->in_stage[i] &&
            hipHostMalloc(&ctx->in_stage[i], kInChunk, hipHostMallocDefault) != hipSuccess) {
            ctx->in_stage[i] = nullptr;
            return fail(PF_E_NOMEM, "hipHostMalloc(%zu) failed", kInChunk);
        }
        if (!ctx->in_ev[i]) PF_HIP(hipEventCreateWithFlags(&ctx->in_ev[i], hipEventDisableTiming));
    }
    return PF_OK;
}

struct Stager {
    pf_ctx* ctx;
    int slot = 0;
    size_t fill = 0;  // bytes in the current chunk
    char* dst = nullptr;
    size_t pos = 0;  // bytes of dst already issued

    void begin(void* d) {
        dst = static_cast<char*>(d);
        pos = 0;
    }
    int wait(int i) {
        if (ctx->in_busy[i]) {
            PF_HIP(hipEventSynchronize(ctx->in_ev[i]));
            ctx->in_busy[i] = false;
        }
        return PF_OK;
    }
    int flush() {
        if (!fill) return PF_OK;
        PF_HIP(hipMemcpyAsync(dst + pos, ctx->in_stage[slot], fill, hipMemcpyHostToDevice,
                              ctx->stream));
        PF_HIP(hipEventRecord(ctx->in_ev[slot], ctx->stream));
        ctx->in_busy[slot] = true;
        pos += fill;
        fill = 0;
        slot ^= 1;
        return wait(slot);
    }
    int put(const void* src, size_t bytes) {
        const char* p = static_cast<const char*>(src);
        while (bytes) {
            const size_t take = bytes < kInChunk - fill ? bytes : kInChunk - fill;
            std::memcpy(static_cast<char*>(ctx->in_stage[slot]) + fill, p, take);
            fill += take;
            p += take;
            bytes -= take;
            if (fill == kInChunk) PF_TRY(flush());
        }
        return PF_OK;
    }
    int drain() {
        PF_TRY(wait(0));
        return wait(1);
    }
};

// Shape of a set from its offsets (host); drops everything derived.
void set_layout(CurveSet& cs, int64_t n_curves, const int64_t* off, int64_t npts, int64_t mx,
                bool has_err) {
    cs.invalidate();
    cs.n_curves = n_curves;
    cs.npts = npts;
    cs.max_len = mx;
    cs.off_h.assign(off, off + n_curves + 1);
    double s1 = 0.0, s2 = 0.0;
    for (int64_t c = 0; c < n_curves; ++c) {
        const double len = static_cast<double>(off[c + 1] - off[c]);
        s1 += len;
        s2 += len * len;
    }
    cs.rep_len = s1 > 0.0 ? static_cast<int64_t>(s2 / s1) : 0;
    cs.has_err = has_err;
}

// Per-curve host arrays (t[i], y[i], err[i] of lengths[i]) into cs.
int upload_lists(pf_ctx* ctx, CurveSet& cs, int64_t n_curves, const float* const* t,
                 const float* const* y, const float* const* e, const int64_t* len) {
    if (n_curves < 0) return fail(PF_E_ARG, "n_curves < 0");
    if (n_curves > 0 && (!len || !t || !y)) return fail(PF_E_ARG, "t, y and lengths must not be NULL");
    std::vector<int64_t> off(static_cast<size_t>(n_curves) + 1, 0);
    int64_t mx = 0;
    for (int64_t c = 0; c < n_curves; ++c) {
        const int64_t L = len[c];
        if (L < 0) return fail(PF_E_ARG, "lengths[%lld] < 0", static_cast<long long>(c));
        if (L > 0 && (!t[c] || !y[c] || (e && !e[c])))
            return fail(PF_E_ARG, "curve %lld: array pointer is NULL", static_cast<long long>(c));
        off[c + 1] = off[c] + L;
        if (L > mx) mx = L;
    }
    const int64_t npts = off[n_curves];
    set_layout(cs, n_curves, off.data(), npts, mx, e != nullptr);
    const size_t nb = static_cast<size_t>(npts) * sizeof(float);
    PF_TRY(ensure(cs.off, (n_curves + 1) * sizeof(int64_t)));
    PF_TRY(ensure(cs.t, nb));
    PF_TRY(ensure(cs.y, nb));
    if (e) PF_TRY(ensure(cs.e, nb));
    PF_TRY(stage_init(ctx));
    Stager st{ctx};
    int rc = PF_OK;
    auto copy = [&](void* dst, const float* const* src) {
        st.begin(dst);
        for (int64_t c = 0; c < n_curves && rc == PF_OK; ++c)
            if (len[c] > 0) rc = st.put(src[c], static_cast<size_t>(len[c]) * sizeof(float));
        if (rc == PF_OK) rc = st.flush();
    };
    st.begin(cs.off.p);
    rc = st.put(off.data(), (n_curves + 1) * sizeof(int64_t));
    if (rc == PF_OK) rc = st.flush();
    if (rc == PF_OK) copy(cs.t.p, t);
    if (rc == PF_OK) copy(cs.y.p, y);
    if (rc == PF_OK && e) copy(cs.e.p, e);
    // The copies complete in stream order, before any later work on the set
    // (every use is on ctx->stream, or on side streams that wait for it);
    // a staging chunk is reused only after its copy's event.  So a load does
    // not wait for its copies, except to leave a failed one quiet.
    if (rc != PF_OK) {
        (void)st.drain();
        set_layout(cs, 0, off.data(), 0, 0, false);  // an empty, consistent set
    }
    return rc;
}

// Copy host CSR arrays into cs (device), invalidating derived arrays.
int upload_curves(pf_ctx* ctx, CurveSet& cs, int64_t n_curves, const int64_t* off,
                  const float* t, const float* y, const float* e) {
    int64_t npts = 0, mx = 0;
    PF_TRY(validate_offsets(n_curves, off, &npts, &mx));
    if (npts > 0 && (!t || !y)) return fail(PF_E_ARG, "t and y must not be NULL");
    set_layout(cs, n_curves, off, npts, mx, e != nullptr);
    const size_t nb = static_cast<size_t>(npts) * sizeof(float);
    const int64_t zero = 0;
    // A failed grow leaves an empty, consistent set.
    auto grow = [&](DevBuf& b, size_t bytes) {
        const int rc = ensure(b, bytes);
        if (rc != PF_OK) set_layout(cs, 0, &zero, 0, 0, false);
        return rc;
    };
    PF_TRY(grow(cs.off, (n_curves + 1) * sizeof(int64_t)));
    PF_TRY(grow(cs.t, nb));
    PF_TRY(grow(cs.y, nb));
    PF_HIP(hipMemcpyAsync(cs.off.p, off, (n_curves + 1) * sizeof(int64_t), hipMemcpyHostToDevice,
                          ctx->stream));
    if (nb) {
        PF_HIP(hipMemcpyAsync(cs.t.p, t, nb, hipMemcpyHostToDevice, ctx->stream));
        PF_HIP(hipMemcpyAsync(cs.y.p, y, nb, hipMemcpyHostToDevice, ctx->stream));
    }
    if (e) {
        PF_TRY(grow(cs.e, nb));
        if (nb) PF_HIP(hipMemcpyAsync(cs.e.p, e, nb, hipMemcpyHostToDevice, ctx->stream));
    }
    return PF_OK;
}

// Matched Filter templates, generate_templates (rust/src/mf.rs:29-112):
// 5 sawtooth (rise fractions), 1 sinusoid, 4 dip widths x {no, half-depth}
// secondary eclipse, each normalised to zero mean and unit variance in f32
// (sums fold from -0.0, as Rust's f32 Sum does; sin is the C library's sinf,
// as Rust std calls it).  Built on the host once per bin count and stored as
// rows of t[j % nb], j < mf_template_stride(nb).
void mf_normalize(std::vector<float>& t) {
    const float n = static_cast<float>(t.size());
    float sum = -0.0f;
    for (float x : t) sum = sum + x;
    const float mean = sum / n;
    for (float& x : t) x -= mean;
    float var = -0.0f;
    for (float x : t) var = var + x * x;
    const float sd = std::sqrt(var / n);
    if (sd > 1e-12f)
        for (float& x : t) x /= sd;
}

std::vector<std::vector<float>> mf_shapes(int nb) {
    std::vector<std::vector<float>> out;
    const float nbf = static_cast<float>(nb);
    auto phase_of = [&](int k) { return (static_cast<float>(k) + 0.5f) / nbf; };
    for (float rf : {0.10f, 0.20f, 0.35f, 0.50f, 0.65f}) {
        std::vector<float> t(nb);
        for (int k = 0; k < nb; ++k) {
            const float ph = phase_of(k);
            t[k] = ph < rf ? ph / rf : 1.0f - (ph - rf) / (1.0f - rf);
        }
        mf_normalize(t);
        out.push_back(std::move(t));
    }
    {
        const float two_pi = 2.0f * static_cast<float>(M_PI);
        std::vector<float> t(nb);
        for (int k = 0; k < nb; ++k) t[k] = sinf(two_pi * phase_of(k));
        mf_normalize(t);
        out.push_back(std::move(t));
    }
    for (float dw : {0.05f, 0.10f, 0.15f, 0.20f}) {
        for (int sec = 0; sec < 2; ++sec) {
            std::vector<float> t(nb, 1.0f);
            const float half = dw / 2.0f;
            for (int k = 0; k < nb; ++k) {
                const float ph = phase_of(k);
                if ((ph < 0.5f ? ph : 1.0f - ph) < half) t[k] = 0.0f;
                if (sec && std::fabs(ph - 0.5f) < half) t[k] = 1.0f - 0.5f;
            }
            mf_normalize(t);
            out.push_back(std::move(t));
        }
    }
    return out;
}

// Device table of the templates for nb bins, built on first use and kept
// (entries are only added once their upload has been issued).
const float* mf_templates(const pf_ctx* ctx, int nb) {
    for (const pf_ctx::MfTable& m : ctx->mf_tables)
        if (m.nb == nb) return static_cast<const float*>(m.buf.p);
    return nullptr;
}

int ensure_mf_templates(pf_ctx* ctx, int nb) {
    if (mf_templates(ctx, nb)) return PF_OK;
    const std::vector<std::vector<float>> sh = mf_shapes(nb);
    const int stride = pf::mf_template_stride(nb);
    pf_ctx::MfTable m;
    m.nb = nb;
    m.h.assign(static_cast<size_t>(pf::kMfShapes) * stride, 0.0f);
    for (int s = 0; s < pf::kMfShapes; ++s)
        for (int j = 0; j < stride; ++j) m.h[s * stride + j] = sh[s][j % nb];
    PF_TRY(ensure(m.buf, m.h.size() * sizeof(float)));
    const hipError_t e = hipMemcpyAsync(m.buf.p, m.h.data(), m.h.size() * sizeof(float),
                                        hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) {
        (void)hipFree(m.buf.p);
        return fail(PF_E_HIP, "MF template upload failed: %s", hipGetErrorString(e));
    }
    ctx->mf_tables.push_back(std::move(m));
    return PF_OK;
}

// Harmonics of MHF's largest model: max_harmonics.min(5) (mhf.rs:238).
int mhf_max_k(int max_harmonics) {
    return max_harmonics < pf::kMhfMaxK ? max_harmonics : pf::kMhfMaxK;
}

// MHF kernel arguments for a plain launch over cs with `tblocks` trial
// blocks per curve.  When the (curve, trial block) pairs are too few to give
// every SIMD several equal work units, each pair's points are cut into
// slices (at least 256 points each) until there are about 8 units per SIMD;
// the slices' partial sums go through a workspace (kept under 1 GiB) and a
// finishing kernel.  Results do not depend on the split but for the order in
// which partial sums are added.
int mhf_args(pf_ctx* ctx, CurveSet& cs, int max_k, int64_t tblocks, float* feat,
             pf::MhfArgs* a) {
    a->pt = as<pf::MhfPoint>(cs.mhf_pt);
    a->sums = as<double>(cs.mhf_sums);
    a->max_k = max_k;
    a->feat = feat;
    a->split = 1;
    a->part = nullptr;
    if (max_k == 0 || std::getenv("PF_MHF_NO_SPLIT")) return PF_OK;
    const int64_t pairs = cs.n_curves * tblocks;
    const int64_t target = 8 * 4 * static_cast<int64_t>(ctx->n_cu);
    if (pairs <= 0 || 2 * pairs >= target) return PF_OK;
    int64_t sp = (target + pairs - 1) / pairs;
    sp = std::min<int64_t>(sp, cs.max_len / 256);
    sp = std::min<int64_t>(sp, 64);
    while (sp > 1 &&
           pf::mhf_part_doubles(static_cast<int>(sp), cs.n_curves, tblocks, max_k) * 8 > (1ull << 30))
        --sp;
    if (sp <= 1) return PF_OK;
    PF_TRY(ensure(cs.mhf_part,
                  pf::mhf_part_doubles(static_cast<int>(sp), cs.n_curves, tblocks, max_k) * 8));
    a->split = static_cast<int>(sp);
    a->part = as<double>(cs.mhf_part);
    return PF_OK;
}

// LS variant of a launch: 1 = hardware revolution sin/cos with the
// uncorrected fold (default), 0 = sincosf of fl(2 pi phase) with the exact
// fold (PF_LS_PRECISE set).  Read at every launch; captured graphs key on it.
int ls_fast_mode() { return std::getenv("PF_LS_PRECISE") ? 0 : 1; }

// The magnitudes a call evaluates: as uploaded, or prepared on the device
// (prepare() must have run for p).
const float* mags(CurveSet& cs, const Params& p) {
    return p.mag_prep == PF_MAG_AS_IS ? as<float>(cs.y) : as<float>(cs.yp);
}

// The per-curve time bound the binned kernels' fast phase bins read
// (prepare() builds it with the f64 times), or nullptr: every bin then
// comes from the exact fold, as with PF_EXACT_BINS set (read at every
// launch; captured graphs key on the pointer).  MHF folds its own records.
const float* time_bound(CurveSet& cs, const Params& p) {
    if (p.algo == PF_ALGO_MHF || std::getenv("PF_EXACT_BINS")) return nullptr;
    return as<float>(cs.tmax);
}

// The packed-count columns of CE and VN (built with the magnitude rows for
// an even n_mag), or nullptr; without them the kernels derive the packed
// cell from the magnitude row.
const uint2* ce_packed_cols(CurveSet& cs, const Params& p) {
    return p.n_mag % 2 == 0 ? as<uint2>(cs.ce_mpk) : nullptr;
}

// Equal weights (every curve's weights bitwise equal, as with constant or
// absent errors): a bin's serial sum of weights is then the curve's running
// sum wsum[c - 1] for its c points, so FPW and BLS count points per bin
// instead of summing weights.  PF_EQUAL_WEIGHTS=0 disables this, =1 makes
// BLS take its counting kernel for every launch it can serve (both for tests
// and comparisons).
int ew_switch() {
    const char* sw = std::getenv("PF_EQUAL_WEIGHTS");
    return sw && sw[0] == '0' ? 0 : (sw && sw[0] == '1' ? 1 : -1);
}

// The running sums and the device flag, once per set (cs.w must be built).
// No wait: the kernels read the flag.  Sets over fixed memory (the
// allocation-free entry point) have neither.
int ensure_wsum(CurveSet& cs, hipStream_t s) {
    if (cs.wsum_ok || cs.wsum.fixed || cs.npts == 0) return PF_OK;
    PF_TRY(ensure(cs.w_flag, sizeof(int)));
    PF_TRY(ensure(cs.wsum, static_cast<size_t>(cs.npts) * sizeof(float)));
    PF_HIP(hipMemsetAsync(cs.w_flag.p, 0, sizeof(int), s));
    if (cs.has_err)  // without errors every weight is 1
        PF_HIP(pf::launch_bls_unequal(as<int64_t>(cs.off), as<float>(cs.w), as<int>(cs.w_flag),
                                      cs.n_curves, s));
    PF_HIP(pf::launch_bls_wsum(as<int64_t>(cs.off), as<float>(cs.w), as<float>(cs.wsum),
                               cs.n_curves, s));
    cs.wsum_ok = true;
    return PF_OK;
}

// BLS's counting kernel (pf::BlsArgs::wsum) fits 8 blocks per CU where the
// general one fits 6, at a somewhat higher cost per block (one more LDS
// operation per point, the weight prefix read from registers in the search).
// So a launch takes it only where that saves a round of blocks, or for the
// persistent launch, and never when it shares the device with other
// statistics' launches (`shared`: overlapped with the jobs of a
// multi-statistic call, the general kernel measured faster for the whole
// call).  Curves beyond the 16-bit counts and more than kBlsEwMaxBins bins
// keep the general kernel.
bool bls_ew_wanted(const pf_ctx* ctx, const CurveSet& cs, const Params& p, int64_t T,
                   bool persistent, bool shared) {
    if (p.algo != PF_ALGO_BLS || p.n_bins > pf::kBlsEwMaxBins) return false;
    if (cs.wsum.fixed || cs.max_len > pf::kBlsEwMaxLen || cs.npts == 0) return false;
    const int sw = ew_switch();
    if (sw >= 0) return sw == 1;
    if (shared) return false;
    if (persistent) return true;
    const int64_t blocks = cs.n_curves * ((T + pf::kWaveHost - 1) / pf::kWaveHost);
    const int64_t slots_g =
        std::max<int64_t>(1, static_cast<int64_t>(ctx->lds_max / pf::lds_bls(p.n_bins))) * ctx->n_cu;
    const int64_t slots_e =
        std::max<int64_t>(1, static_cast<int64_t>(ctx->lds_max / pf::lds_bls_ew(p.n_bins))) *
        ctx->n_cu;
    return (blocks + slots_g - 1) / slots_g > (blocks + slots_e - 1) / slots_e;
}

// FPW counts points per bin whenever the weights are equal: the count takes
// the place of the weight plane, the kernel reads the device flag itself.
bool fpw_ew_wanted(const CurveSet& cs, const Params& p) {
    return p.algo == PF_ALGO_FPW && !cs.wsum.fixed && cs.npts > 0 && ew_switch() != 0;
}

// The equal-weight inputs a launch of `p` will use, before it is queued (and
// outside graph capture: BLS reads the flag back, which waits for the
// stream, once per set).
int prepare_equal_weights(const pf_ctx* ctx, CurveSet& cs, const Params& p, int64_t T,
                          bool persistent, bool shared, hipStream_t s) {
    if (fpw_ew_wanted(cs, p)) return ensure_wsum(cs, s);
    if (!bls_ew_wanted(ctx, cs, p, T, persistent, shared)) return PF_OK;
    PF_TRY(ensure_wsum(cs, s));
    if (!cs.w_equal_known) {
        int unequal = 0;
        PF_HIP(hipMemcpyAsync(&unequal, cs.w_flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
        PF_HIP(hipStreamSynchronize(s));
        cs.w_equal = unequal == 0;
        cs.w_equal_known = true;
    }
    return PF_OK;
}

// Build (or reuse) the derived arrays params.algo needs.
int prepare(pf_ctx* ctx, CurveSet& cs, const Params& p, bool has_pdt, hipStream_t s) {
    const int64_t n = cs.npts;
    const float* e = cs.has_err ? as<float>(cs.e) : nullptr;
    const int64_t* off = as<int64_t>(cs.off);
    // Magnitudes in the requested preparation; what was derived from another
    // preparation is rebuilt.
    // Each cached array's validity is cleared before its buffer may be
    // regrown: a failed grow frees the old buffer.
    if (p.mag_prep != PF_MAG_AS_IS && cs.yp_mode != p.mag_prep) {
        cs.yp_mode = -1;
        PF_TRY(ensure(cs.yp, n * sizeof(float)));
        PF_HIP(pf::launch_prep_mags(off, as<float>(cs.y), as<float>(cs.yp), cs.n_curves,
                                    p.mag_prep, s));
        cs.yp_mode = p.mag_prep;
    }
    if (cs.derived_mode != p.mag_prep) {
        cs.invalidate_mags();
        cs.derived_mode = p.mag_prep;
    }
    const float* y = mags(cs, p);
    // (MHF reads times from its own point records.)
    if (p.algo != PF_ALGO_MHF && (!cs.t64_ok || (has_pdt && !cs.tt_ok))) {
        cs.t64_ok = cs.tt_ok = false;
        PF_TRY(ensure(cs.t64, n * sizeof(double)));
        if (has_pdt) PF_TRY(ensure(cs.tt, n * sizeof(double)));
        PF_TRY(ensure(cs.tmax, static_cast<size_t>(cs.n_curves) * sizeof(float)));
        PF_HIP(pf::launch_widen_times(as<float>(cs.t), as<double>(cs.t64),
                                      has_pdt ? as<double>(cs.tt) : nullptr, n, s));
        PF_HIP(pf::launch_time_bound(off, as<float>(cs.t), as<float>(cs.tmax), cs.n_curves, s));
        cs.t64_ok = true;
        cs.tt_ok = has_pdt;
    }
    switch (p.algo) {
        case PF_ALGO_VN:
            if (cs.vn_ext != p.phase_ext * p.mag_ext) {
                const int ext = p.phase_ext * p.mag_ext;
                cs.vn_ext = -1;
                std::vector<int64_t>& lo = cs.vn_lut_off_h;
                lo.assign(static_cast<size_t>(cs.n_curves), 0);
                int64_t total = 0;
                for (int64_t c = 0; c < cs.n_curves; ++c) {
                    lo[c] = total;
                    total += (cs.off_h[c + 1] - cs.off_h[c]) * ext + 1;
                }
                PF_TRY(ensure(cs.vn_lut_off, lo.size() * sizeof(int64_t)));
                PF_TRY(ensure(cs.vn_lut, static_cast<size_t>(total) * sizeof(float)));
                if (!lo.empty())
                    PF_HIP(hipMemcpyAsync(cs.vn_lut_off.p, lo.data(), lo.size() * sizeof(int64_t),
                                          hipMemcpyHostToDevice, s));
                PF_HIP(pf::launch_vn_lut(off, as<int64_t>(cs.vn_lut_off), as<float>(cs.vn_lut),
                                         cs.n_curves, ext, s));
                cs.vn_ext = ext;
            }
            [[fallthrough]];
        case PF_ALGO_CE:
            if (cs.ce_nmag != p.n_mag) {
                cs.ce_nmag = -1;
                const bool even = p.n_mag % 2 == 0;
                PF_TRY(ensure(cs.ce_mrow, n * sizeof(uint32_t)));
                if (even) PF_TRY(ensure(cs.ce_mpk, n * sizeof(uint2)));
                PF_HIP(pf::launch_ce_mag_rows(y, as<uint32_t>(cs.ce_mrow),
                                              even ? as<uint2>(cs.ce_mpk) : nullptr, n,
                                              p.n_mag, s));
                cs.ce_nmag = p.n_mag;
            }
            break;
        case PF_ALGO_AOV:
            if (!cs.aov_ok) {
                PF_TRY(ensure(cs.aov_yy, n * sizeof(float)));
                PF_TRY(ensure(cs.aov_avg, cs.n_curves * sizeof(float)));
                PF_HIP(pf::launch_square(y, as<float>(cs.aov_yy), n, s));
                PF_HIP(pf::launch_curve_mean(off, y, as<float>(cs.aov_avg), cs.n_curves, s));
                cs.aov_ok = true;
            }
            break;
        case PF_ALGO_FPW:
        case PF_ALGO_BLS:
        case PF_ALGO_MF:
            if (!cs.w_ok) {
                PF_TRY(ensure(cs.w, n * sizeof(float)));
                PF_HIP(pf::launch_inv_var(e, as<float>(cs.w), n, s));
                cs.w_ok = true;
            }
            if ((p.algo == PF_ALGO_FPW || p.algo == PF_ALGO_MF) && !cs.fpw_ok) {
                PF_TRY(ensure(cs.fpw_wy, n * sizeof(float)));
                PF_HIP(pf::launch_mul(as<float>(cs.w), y, as<float>(cs.fpw_wy), n, s));
                cs.fpw_ok = true;
            }
            if (p.algo == PF_ALGO_MF && !cs.mf_ok) {  // (w*y)*y, mf.rs:187
                PF_TRY(ensure(cs.mf_wyy, n * sizeof(float)));
                PF_HIP(pf::launch_mul(as<float>(cs.fpw_wy), y, as<float>(cs.mf_wyy), n, s));
                cs.mf_ok = true;
            }
            if (p.algo == PF_ALGO_MF) PF_TRY(ensure_mf_templates(ctx, p.n_bins));
            if (p.algo == PF_ALGO_BLS && !cs.bls_ok) {
                PF_TRY(ensure(cs.bls_yw, n * sizeof(float)));
                PF_TRY(ensure(cs.bls_mean, cs.n_curves * sizeof(float)));
                PF_HIP(pf::launch_bls_prepare(off, y, as<float>(cs.w), as<float>(cs.bls_mean),
                                              as<float>(cs.bls_yw), cs.n_curves, cs.max_len, s));
                cs.bls_ok = true;
            }
            break;
        case PF_ALGO_MHF:
            if (!cs.mhf_ok) {
                PF_TRY(ensure(cs.mhf_pt, n * sizeof(pf::MhfPoint)));
                PF_TRY(ensure(cs.mhf_sums, cs.n_curves * pf::kMhfSums * sizeof(double)));
                PF_HIP(pf::launch_mhf_prep(off, as<float>(cs.t), y, e, as<pf::MhfPoint>(cs.mhf_pt),
                                           as<double>(cs.mhf_sums), cs.n_curves, s));
                cs.mhf_ok = true;
            }
            break;
        default:
            break;
    }
    return PF_OK;
}

// The trial grid a launch reads: device periods[np], period_dts[nd].
struct TrialGrid {
    const float* periods;
    const float* pdts;
    bool has_pdt;  // any period_dt != 0
};

TrialGrid ctx_grid(const pf_ctx* ctx) {
    return TrialGrid{static_cast<const float*>(ctx->periods.p),
                     static_cast<const float*>(ctx->pdts.p), ctx->has_pdt};
}

// Launch the grid kernel of params.algo over cs into cs.out ([n_curves][T])
// on stream s.  With q, the persistent kernel pulls work from *q (derived
// arrays must already be prepared).  `timed` brackets the launch with the
// context's events (not allowed while the stream is being captured).
int launch_grid(pf_ctx* ctx, CurveSet& cs, const Params& p, int64_t np, int64_t nd,
                const TrialGrid& tg, const pf::WorkQueue* q, bool timed, int group,
                hipStream_t s, bool shared = false) {
    const int64_t T = np * nd;
    const bool has_pdt = tg.has_pdt;

    pf::GridArgs g;
    g.off = as<int64_t>(cs.off);
    g.t64 = as<double>(cs.t64);
    g.tt = has_pdt ? as<double>(cs.tt) : as<double>(cs.t64);
    g.periods = tg.periods;
    g.pdts = tg.pdts;
    g.nd = nd;
    g.T = T;
    g.tblocks = (T + pf::kWaveHost - 1) / pf::kWaveHost;
    g.n_blocks = static_cast<uint64_t>(cs.n_curves) * static_cast<uint64_t>(g.tblocks);
    g.out = as<float>(cs.out);
    g.has_pdt = has_pdt ? 1 : 0;
    g.tmax = time_bound(cs, p);
    if (g.n_blocks > 0xffffffffull)
        return fail(PF_E_LIMIT, "grid of %llu workgroups exceeds one launch",
                    static_cast<unsigned long long>(g.n_blocks));

    // MHF's launch shape (and any workspace) is settled before the timed region.
    pf::MhfArgs mhf{as<pf::MhfPoint>(cs.mhf_pt), as<double>(cs.mhf_sums),
                    mhf_max_k(p.max_harmonics), nullptr, 1, nullptr};
    if (p.algo == PF_ALGO_MHF && !q)  // persistent (possibly captured): one kernel, no workspace
        PF_TRY(mhf_args(ctx, cs, mhf_max_k(p.max_harmonics), g.tblocks, nullptr, &mhf));

    if (timed) PF_HIP(hipEventRecord(ctx->ev0, s));
    hipError_t le = hipSuccess;
    switch (p.algo) {
        case PF_ALGO_CE: {
            pf::CeArgs a{as<uint32_t>(cs.ce_mrow), p.n_phase, p.n_mag, p.phase_ext, p.mag_ext,
                         ce_packed_cols(cs, p)};
            // The atomic variant counts in 16-bit halves (half the LDS, twice
            // the waves per CU) whenever no count can exceed 65535.
            const bool can_pack = p.phase_ext == 1 && p.mag_ext == 1 &&
                                  cs.max_len <= pf::kCePackMaxLen;
            const size_t lds_full = pf::lds_ce(p.n_phase, p.n_mag);
            const size_t lds_pack = pf::lds_ce_packed(p.n_phase, p.n_mag);
            if (q) {
                pf::WorkQueue qq = *q;
                bool all_packed = true;
                for (int b = 0; b < qq.n_buckets; ++b) {
                    if (qq.bucket_group[b] == 0 && can_pack) qq.bucket_group[b] = -1;
                    if (qq.bucket_group[b] >= 0) all_packed = false;
                }
                le = pf::launch_ce(g, a, all_packed ? lds_pack : lds_full, all_packed ? -1 : 0,
                                   &qq, s);
            } else {
                const int grp = (group == 0 && can_pack) ? -1 : group;
                le = pf::launch_ce(g, a, grp < 0 ? lds_pack : lds_full, grp, nullptr, s);
            }
            break;
        }
        case PF_ALGO_AOV: {
            pf::AovArgs a{mags(cs, p), as<float>(cs.aov_yy), as<float>(cs.aov_avg),
                          p.n_phase, p.phase_ext};
            le = pf::launch_aov(g, a, pf::lds_aov(p.n_phase), group, q, s);
            break;
        }
        case PF_ALGO_LS: {
            pf::LsArgs a{mags(cs, p), ls_fast_mode()};
            le = pf::launch_ls(g, a, q, s);
            break;
        }
        case PF_ALGO_FPW: {
            const bool ew = cs.wsum_ok && fpw_ew_wanted(cs, p);
            pf::FpwArgs a{as<float>(cs.w), as<float>(cs.fpw_wy), p.n_bins,
                          ew ? as<float>(cs.wsum) : nullptr, ew ? as<int>(cs.w_flag) : nullptr};
            le = pf::launch_fpw(g, a, pf::lds_fpw(p.n_bins), group, q, s);
            break;
        }
        case PF_ALGO_VN: {
            pf::VnArgs a{{as<uint32_t>(cs.ce_mrow), p.n_phase, p.n_mag, p.phase_ext, p.mag_ext,
                          ce_packed_cols(cs, p)},
                         p.bandwidth, p.margin, as<float>(cs.vn_lut), as<int64_t>(cs.vn_lut_off)};
            // 16-bit packed counts when no count can exceed 65535 and every row
            // owns a whole word for the trace's per-row fraction (n_mag >= 3).
            const bool can_pack = p.phase_ext == 1 && p.mag_ext == 1 && cs.max_len <= 65535 &&
                                  p.n_mag >= 3;
            const size_t lds = can_pack ? pf::lds_ce_packed(p.n_phase, p.n_mag)
                                        : pf::lds_ce(p.n_phase, p.n_mag);
            if (q) {
                pf::WorkQueue qq = *q;
                for (int b = 0; b < qq.n_buckets; ++b) qq.bucket_group[b] = can_pack ? -1 : 0;
                le = pf::launch_vn(g, a, lds, 0, &qq, s);
            } else {
                le = pf::launch_vn(g, a, lds, can_pack ? -1 : 0, nullptr, s);
            }
            break;
        }
        case PF_ALGO_MF: {
            pf::MfArgs a{as<float>(cs.w), as<float>(cs.fpw_wy), as<float>(cs.mf_wyy), mags(cs, p),
                         mf_templates(ctx, p.n_bins), p.n_bins, nullptr, nullptr};
            le = pf::launch_mf(g, a, pf::lds_mf(p.n_bins), group, q, s);
            break;
        }
        case PF_ALGO_MHF:
            le = pf::launch_mhf(g, mhf, q, s);
            break;
        case PF_ALGO_BLS: {
            int lo, hi;
            bls_duration_range(p.n_bins, p.qmin, p.qmax, &lo, &hi);
            // Equal weights: the counting kernel where bls_ew_wanted (the
            // caller ran prepare_equal_weights).
            const bool ew = cs.w_equal_known && cs.w_equal &&
                            bls_ew_wanted(ctx, cs, p, T, q != nullptr, shared);
            pf::BlsArgs a{as<float>(cs.w), as<float>(cs.bls_yw), p.n_bins, lo, hi,
                          ew ? as<float>(cs.wsum) : nullptr};
            le = pf::launch_bls(g, a, ew ? pf::lds_bls_ew(p.n_bins) : pf::lds_bls(p.n_bins),
                                group, q, s);
            break;
        }
    }
    if (le != hipSuccess)
        return fail(PF_E_HIP, "grid kernel launch failed: %s", hipGetErrorString(le));
    if (timed) {
        PF_HIP(hipEventRecord(ctx->ev1, s));
        ctx->timed = true;
    }
    return PF_OK;
}

int run_grid(pf_ctx* ctx, CurveSet& cs, const Params& p, int64_t np, int64_t nd) {
    PF_TRY(check_params(p, ctx->lds_max));
    const int64_t T = np * nd;
    ctx->timed = false;
    if (cs.n_curves == 0 || T == 0) return PF_OK;
    PF_TRY(prepare(ctx, cs, p, ctx->has_pdt, ctx->stream));
    PF_TRY(prepare_equal_weights(ctx, cs, p, T, false, false, ctx->stream));
    cs.out_T = 0;
    PF_TRY(ensure(cs.out, static_cast<size_t>(cs.n_curves) * T * sizeof(float)));
    cs.out_T = T;
    return launch_grid(ctx, cs, p, np, nd, ctx_grid(ctx), nullptr, true,
                       tune_choice(ctx, p.algo, cs.rep_len).group, ctx->stream);
}

int finish(pf_ctx* ctx) {
    PF_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->timed) {
        float ms = 0.0f;
        PF_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        ctx->last_ms = ms;
    }
    return PF_OK;
}

// The outputs of one call live back to back (16-byte aligned) in
// ctx->outbuf; one device-to-host copy brings them into pinned staging, and
// they are then copied to the caller's arrays.  One transfer per call instead
// of one per output array.
struct OutRegion {
    void* host;
    size_t bytes;
    size_t off;
};

struct OutPlan {
    std::vector<OutRegion> r;
    size_t total = 0;
    size_t add(void* host, size_t bytes) {
        const size_t off = (total + 15) & ~size_t(15);
        r.push_back(OutRegion{host, bytes, off});
        total = off + bytes;
        return off;
    }
};

template <typename T>
T* out_dev(pf_ctx* ctx, size_t off) {
    return reinterpret_cast<T*>(static_cast<char*>(ctx->outbuf.p) + off);
}

// Workspace of a peak scan over n_curves x T, grown as the split needs it
// (the split reads it only when it splits; see pf::peaks_segments).
int peaks_workspace(pf_ctx* ctx, int64_t n_curves, int64_t T, int64_t n_peaks,
                    int64_t min_distance, pf::PeaksWorkspace* ws) {
    const int n_seg = pf::peaks_segments(n_curves, T, min_distance);
    const size_t n = static_cast<size_t>(n_curves) * n_seg * static_cast<size_t>(n_peaks);
    const size_t idx_bytes = (n * sizeof(int64_t) + 255) & ~size_t(255);
    PF_TRY(ensure(ctx->peaks_ws, idx_bytes + n * sizeof(float)));
    ws->idx = as<int64_t>(ctx->peaks_ws);
    ws->val = reinterpret_cast<float*>(static_cast<char*>(ctx->peaks_ws.p) + idx_bytes);
    return PF_OK;
}

int ensure_host(pf_ctx* ctx, size_t bytes) {
    if (ctx->host_cap >= bytes) return PF_OK;
    if (ctx->host_stage) {
        (void)hipHostFree(ctx->host_stage);
        ctx->host_stage = nullptr;
        ctx->host_cap = 0;
    }
    const size_t want = bytes + bytes / 4 + 4096;
    if (hipHostMalloc(&ctx->host_stage, want, hipHostMallocDefault) != hipSuccess) {
        ctx->host_stage = nullptr;
        return fail(PF_E_NOMEM, "hipHostMalloc(%zu) failed", want);
    }
    ctx->host_cap = want;
    return PF_OK;
}

// Queue the copy of every region, wait for the stream, scatter to the caller.
int out_finish(pf_ctx* ctx, const OutPlan& plan) {
    PF_TRY(ensure_host(ctx, plan.total));
    if (plan.total)
        PF_HIP(hipMemcpyAsync(ctx->host_stage, ctx->outbuf.p, plan.total, hipMemcpyDeviceToHost,
                              ctx->stream));
    PF_TRY(finish(ctx));
    for (const OutRegion& o : plan.r)
        if (o.bytes) std::memcpy(o.host, static_cast<char*>(ctx->host_stage) + o.off, o.bytes);
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
    q.mag_prep = p->mag_prep;
    q.bandwidth = p->bandwidth;
    q.margin = p->margin;
    q.max_harmonics = p->max_harmonics;
    return q;
}

// Grid (+ optional copy-out) over a set already on the device.
int grid_on(pf_ctx* ctx, CurveSet& cs, const Params& p, const float* periods, int64_t np,
            const float* pdts, int64_t nd, float* out) {
    PF_TRY(check_params(p, ctx->lds_max));
    PF_TRY(check_grid(periods, np, pdts, nd));
    if (cs.n_curves > 0 && np * nd > 0) PF_TRY(upload_grid(ctx, periods, np, pdts, nd));
    PF_TRY(run_grid(ctx, cs, p, np, nd));
    const size_t ob = static_cast<size_t>(cs.n_curves) * np * nd * sizeof(float);
    if (out && ob)
        PF_HIP(hipMemcpyAsync(out, cs.out.p, ob, hipMemcpyDeviceToHost, ctx->stream));
    return finish(ctx);
}

// Peaks over a set already on the device.
int peaks_on(pf_ctx* ctx, CurveSet& cs, const Params& p, const float* periods, int64_t np,
             const float* pdts, int64_t nd, int64_t n_peaks, int64_t min_distance, int64_t* idx,
             float* val) {
    PF_TRY(check_params(p, ctx->lds_max));
    PF_TRY(check_grid(periods, np, pdts, nd));
    if (n_peaks < 0 || n_peaks > 64) return fail(PF_E_LIMIT, "n_peaks must be in [0, 64]");
    if (min_distance < 0) return fail(PF_E_ARG, "min_distance < 0");
    if (cs.n_curves > 0 && n_peaks > 0 && (!idx || !val)) return fail(PF_E_ARG, "idx/val is NULL");
    const size_t pk = static_cast<size_t>(cs.n_curves) * n_peaks;
    if (pk == 0) return PF_OK;
    const int64_t T = np * nd;
    if (T == 0) {
        for (size_t i = 0; i < pk; ++i) {
            idx[i] = -1;
            val[i] = 0.0f;
        }
        return PF_OK;
    }
    PF_TRY(upload_grid(ctx, periods, np, pdts, nd));
    PF_TRY(run_grid(ctx, cs, p, np, nd));
    OutPlan plan;
    const size_t o_idx = plan.add(idx, pk * sizeof(int64_t));
    const size_t o_val = plan.add(val, pk * sizeof(float));
    PF_TRY(ensure(ctx->outbuf, plan.total));
    pf::PeaksWorkspace ws;
    PF_TRY(peaks_workspace(ctx, cs.n_curves, T, n_peaks, min_distance, &ws));
    const int use_max = p.algo == PF_ALGO_CE ? 0 : 1;
    PF_HIP(pf::launch_peaks(as<float>(cs.out), cs.n_curves, T, n_peaks, min_distance, use_max,
                            out_dev<int64_t>(ctx, o_idx), out_dev<float>(ctx, o_val),
                            ctx->stream, &ws));
    return out_finish(ctx, plan);
}

// 'stats' over a set already on the device.
int stats_on(pf_ctx* ctx, CurveSet& cs, const Params& p, const float* periods, int64_t np,
             const float* pdts, int64_t nd, int n_stats, float* summary, int64_t* idx,
             float* val, int* fallback) {
    PF_TRY(check_params(p, ctx->lds_max));
    PF_TRY(check_grid(periods, np, pdts, nd));
    const int64_t T = np * nd;
    if (cs.n_curves == 0) return PF_OK;
    if (T < 1 || T > pf::kStatsMaxT)
        return fail(PF_E_LIMIT, "device stats need 1 <= n_periods*n_pdts <= %lld",
                    static_cast<long long>(pf::kStatsMaxT));
    if (n_stats < 1 || n_stats > T) return fail(PF_E_ARG, "n_stats must be in [1, T]");
    if (!summary || !idx || !val || !fallback) return fail(PF_E_ARG, "NULL output");
    if (pf::stats_lds_bytes(T) + pf::stats_lds_static_bytes() > ctx->lds_max)
        return fail(PF_E_LIMIT, "stats exceed LDS");
    PF_TRY(upload_grid(ctx, periods, np, pdts, nd));
    PF_TRY(run_grid(ctx, cs, p, np, nd));
    const size_t C = static_cast<size_t>(cs.n_curves);
    OutPlan plan;
    const size_t o_idx = plan.add(idx, C * n_stats * sizeof(int64_t));
    const size_t o_val = plan.add(val, C * n_stats * sizeof(float));
    const size_t o_sum = plan.add(summary, C * 4 * sizeof(float));
    const size_t o_flg = plan.add(fallback, C * sizeof(int));
    PF_TRY(ensure(ctx->outbuf, plan.total));
    const int use_max = p.algo == PF_ALGO_CE ? 0 : 1;
    PF_HIP(pf::launch_stats(as<float>(cs.out), cs.n_curves, T, n_stats, use_max,
                            out_dev<float>(ctx, o_sum), out_dev<int64_t>(ctx, o_idx),
                            out_dev<float>(ctx, o_val), out_dev<int>(ctx, o_flg), ctx->stream));
    return out_finish(ctx, plan);
}

// Rank -> curve index, longest curve first (stable), uploaded once per set.
int ensure_order(pf_ctx* ctx, CurveSet& cs) {
    if (cs.order_ok) return PF_OK;
    std::vector<int64_t> ord(cs.n_curves);
    for (int64_t c = 0; c < cs.n_curves; ++c) ord[c] = c;
    std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) {
        return (cs.off_h[a + 1] - cs.off_h[a]) > (cs.off_h[b + 1] - cs.off_h[b]);
    });
    PF_TRY(ensure(cs.order, cs.n_curves * sizeof(int64_t)));
    PF_TRY(ensure(cs.counter, kMaxTuneBuckets * sizeof(unsigned long long)));
    PF_HIP(hipMemcpyAsync(cs.order.p, ord.data(), cs.n_curves * sizeof(int64_t),
                          hipMemcpyHostToDevice, ctx->stream));
    PF_HIP(hipStreamSynchronize(ctx->stream));
    cs.order_h.swap(ord);
    cs.order_ok = true;
    return PF_OK;
}

// Ragged peaks: one persistent launch over every (curve, trial block) item,
// longest curves first, then the peak kernel.  The sequence (counter reset,
// grid, peaks) is captured once as a hipGraph and replayed while the inputs,
// buffers and parameters it was captured with are unchanged.
int ragged_peaks_on(pf_ctx* ctx, CurveSet& cs, const Params& p, const float* periods,
                    int64_t np, const float* pdts, int64_t nd, int64_t n_peaks,
                    int64_t min_distance, int64_t* idx, float* val) {
    PF_TRY(check_params(p, ctx->lds_max));
    PF_TRY(check_grid(periods, np, pdts, nd));
    if (n_peaks < 0 || n_peaks > 64) return fail(PF_E_LIMIT, "n_peaks must be in [0, 64]");
    if (min_distance < 0) return fail(PF_E_ARG, "min_distance < 0");
    if (cs.n_curves > 0 && n_peaks > 0 && (!idx || !val)) return fail(PF_E_ARG, "idx/val is NULL");
    const size_t pk = static_cast<size_t>(cs.n_curves) * n_peaks;
    if (pk == 0) return PF_OK;
    const int64_t T = np * nd;
    if (T == 0) {
        for (size_t i = 0; i < pk; ++i) {
            idx[i] = -1;
            val[i] = 0.0f;
        }
        return PF_OK;
    }
    PF_TRY(upload_grid(ctx, periods, np, pdts, nd));
    PF_TRY(prepare(ctx, cs, p, ctx->has_pdt, ctx->stream));
    PF_TRY(prepare_equal_weights(ctx, cs, p, T, true, false, ctx->stream));
    cs.out_T = 0;
    PF_TRY(ensure(cs.out, static_cast<size_t>(cs.n_curves) * T * sizeof(float)));
    cs.out_T = T;
    OutPlan plan;
    const size_t o_idx = plan.add(idx, pk * sizeof(int64_t));
    const size_t o_val = plan.add(val, pk * sizeof(float));
    PF_TRY(ensure(ctx->outbuf, plan.total));
    pf::PeaksWorkspace ws;
    PF_TRY(peaks_workspace(ctx, cs.n_curves, T, n_peaks, min_distance, &ws));
    PF_TRY(ensure_order(ctx, cs));
    const int64_t tblocks = (T + pf::kWaveHost - 1) / pf::kWaveHost;
    // Length buckets of the tuning table over the longest-first order: each
    // is a contiguous rank range with its own binning variant and claim
    // bound, all served by one persistent launch (one drain tail).
    struct Bucket {
        int64_t r0, r1;
        Choice ch;
    };
    std::vector<Bucket> buckets;
    {
        const std::vector<int64_t> edges = tune_edges(ctx, p.algo);
        int64_t r = 0;
        for (size_t b = edges.size(); b-- > 0;) {
            const int64_t r0 = r;
            while (r < cs.n_curves &&
                   cs.off_h[cs.order_h[r] + 1] - cs.off_h[cs.order_h[r]] >= edges[b])
                ++r;
            if (r > r0) buckets.push_back(Bucket{r0, r, tune_choice(ctx, p.algo, edges[b])});
        }
        // Adjacent buckets with the same choices are one bucket; past the
        // queue's capacity, the shortest buckets join the last one kept.
        std::vector<Bucket> merged;
        for (const Bucket& b : buckets) {
            if (!merged.empty() && merged.back().ch.group == b.ch.group &&
                merged.back().ch.claim == b.ch.claim)
                merged.back().r1 = b.r1;
            else if (static_cast<int>(merged.size()) == pf::kQueueBuckets)
                merged.back().r1 = b.r1;
            else
                merged.push_back(b);
        }
        buckets.swap(merged);
    }
    const int use_max = p.algo == PF_ALGO_CE ? 0 : 1;
    // Everything the captured kernels read as arguments.
    std::vector<uint64_t> key = {
        static_cast<uint64_t>(p.algo), static_cast<uint64_t>(p.n_phase),
        static_cast<uint64_t>(p.n_mag), static_cast<uint64_t>(p.phase_ext),
        static_cast<uint64_t>(p.mag_ext), static_cast<uint64_t>(p.n_bins),
        static_cast<uint64_t>(__builtin_bit_cast(uint32_t, p.qmin)),
        static_cast<uint64_t>(__builtin_bit_cast(uint32_t, p.qmax)),
        static_cast<uint64_t>(p.mag_prep), reinterpret_cast<uint64_t>(mags(cs, p)),
        static_cast<uint64_t>(np), static_cast<uint64_t>(nd), static_cast<uint64_t>(n_peaks),
        static_cast<uint64_t>(min_distance), static_cast<uint64_t>(ctx->has_pdt),
        reinterpret_cast<uint64_t>(ctx->periods.p), reinterpret_cast<uint64_t>(ctx->pdts.p),
        reinterpret_cast<uint64_t>(ctx->outbuf.p), static_cast<uint64_t>(o_val),
        reinterpret_cast<uint64_t>(cs.out.p), reinterpret_cast<uint64_t>(cs.t64.p),
        reinterpret_cast<uint64_t>(cs.tt.p), reinterpret_cast<uint64_t>(cs.ce_mrow.p),
        reinterpret_cast<uint64_t>(cs.ce_mpk.p),
        reinterpret_cast<uint64_t>(cs.aov_yy.p), reinterpret_cast<uint64_t>(cs.w.p),
        reinterpret_cast<uint64_t>(cs.fpw_wy.p), reinterpret_cast<uint64_t>(cs.bls_yw.p),
        reinterpret_cast<uint64_t>(cs.wsum_ok ? cs.wsum.p : nullptr),
        static_cast<uint64_t>(cs.w_equal_known && cs.w_equal),
        reinterpret_cast<uint64_t>(cs.order.p), reinterpret_cast<uint64_t>(cs.counter.p),
        reinterpret_cast<uint64_t>(cs.mf_wyy.p),
        reinterpret_cast<uint64_t>(p.algo == PF_ALGO_MF ? mf_templates(ctx, p.n_bins) : nullptr),
        reinterpret_cast<uint64_t>(cs.vn_lut.p), reinterpret_cast<uint64_t>(cs.vn_lut_off.p),
        static_cast<uint64_t>(p.bandwidth), static_cast<uint64_t>(p.margin),
        reinterpret_cast<uint64_t>(cs.mhf_pt.p), reinterpret_cast<uint64_t>(cs.mhf_sums.p),
        static_cast<uint64_t>(mhf_max_k(p.max_harmonics)),
        static_cast<uint64_t>(p.algo == PF_ALGO_LS ? ls_fast_mode() : 0),
        reinterpret_cast<uint64_t>(time_bound(cs, p)),
        reinterpret_cast<uint64_t>(ws.idx), reinterpret_cast<uint64_t>(ws.val)};
    for (const Bucket& b : buckets) {
        key.push_back(static_cast<uint64_t>(b.r0));
        key.push_back(static_cast<uint64_t>(b.r1));
        key.push_back(static_cast<uint64_t>(b.ch.group));
        key.push_back(static_cast<uint64_t>(b.ch.claim));
    }
    if (!cs.graph || cs.graph_key != key) {
        cs.drop_graph();
        hipStream_t s = ctx->stream;
        unsigned long long* counters = as<unsigned long long>(cs.counter);
        PF_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int rc = PF_OK;
        // one counter per queue part (XCD)
        if (pf::launch_zero_u64(counters, pf::kQueuePartsHost, s) != hipSuccess)
            rc = fail(PF_E_HIP, "counter reset launch failed during capture");
        if (rc == PF_OK) {
            pf::WorkQueue q{};
            q.counter = counters;
            q.order = as<int64_t>(cs.order);
            q.total = cs.n_curves * tblocks;
            q.n_buckets = static_cast<int>(buckets.size());
            for (size_t b = 0; b < buckets.size(); ++b) {
                q.bucket_group[b] = buckets[b].ch.group;
                q.bucket_claim[b] = buckets[b].ch.claim;
                q.bucket_end[b] = buckets[b].r1;
            }
            rc = launch_grid(ctx, cs, p, np, nd, ctx_grid(ctx), &q, false, buckets[0].ch.group,
                             s);
        }
        if (rc == PF_OK &&
            pf::launch_peaks(as<float>(cs.out), cs.n_curves, T, n_peaks, min_distance, use_max,
                             out_dev<int64_t>(ctx, o_idx), out_dev<float>(ctx, o_val), s,
                             &ws) != hipSuccess)
            rc = fail(PF_E_HIP, "peaks launch failed during capture");
        hipGraph_t graph = nullptr;
        hipError_t ec = hipStreamEndCapture(s, &graph);
        if (rc != PF_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (ec != hipSuccess) return fail(PF_E_HIP, "graph capture failed: %s", hipGetErrorString(ec));
        hipError_t ei = hipGraphInstantiate(&cs.graph, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            cs.graph = nullptr;
            return fail(PF_E_HIP, "graph instantiate failed: %s", hipGetErrorString(ei));
        }
        cs.graph_key = key;
    }
    PF_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    PF_HIP(hipGraphLaunch(cs.graph, ctx->stream));
    PF_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    ctx->timed = true;
    return out_finish(ctx, plan);
}

// Host in -> host out (the drop-in calls): upload into the context's set.
int grid_host(pf_ctx* ctx, const Params& p, int64_t n_curves, const int64_t* off, const float* t,
              const float* y, const float* e, const float* periods, int64_t np,
              const float* pdts, int64_t nd, float* out) {
    if (!ctx) return fail(PF_E_ARG, "ctx is NULL");
    PF_TRY(check_params(p, ctx->lds_max));
    PF_TRY(check_grid(periods, np, pdts, nd));
    if (n_curves > 0 && np * nd > 0 && !out) return fail(PF_E_ARG, "out is NULL");
    PF_HIP(hipSetDevice(ctx->device));
    PF_TRY(upload_curves(ctx, ctx->scratch, n_curves, off, t, y, e));
    return grid_on(ctx, ctx->scratch, p, periods, np, pdts, nd, out);
}

// --- Differentiable Lomb-Scargle -------------------------------------------

// A device-pointer argument must be device memory of ctx's device.
int check_dev(pf_ctx* ctx, const void* ptr, const char* name) {
    if (!ptr) return fail(PF_E_ARG, "%s is NULL", name);
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, ptr) != hipSuccess) {
        (void)hipGetLastError();
        return fail(PF_E_ARG, "%s is not HIP device memory", name);
    }
    if (a.type != hipMemoryTypeDevice && !a.isManaged)
        return fail(PF_E_ARG, "%s is not HIP device memory", name);
    if (a.device != ctx->device)
        return fail(PF_E_ARG, "%s lives on device %d, the context on device %d", name, a.device,
                    ctx->device);
    return PF_OK;
}

int check_dev_opt(pf_ctx* ctx, const void* ptr, const char* name) {
    return ptr ? check_dev(ctx, ptr, name) : PF_OK;
}

// A device array of `bytes` bytes at ptr must lie inside one HIP allocation
// of ctx's device (a caller's sub-allocation is checked against the whole
// allocation it sits in).
int check_dev_span(pf_ctx* ctx, const void* ptr, size_t bytes, const char* name) {
    PF_TRY(check_dev(ctx, ptr, name));
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, const_cast<void*>(ptr)) != hipSuccess) {
        (void)hipGetLastError();
        return fail(PF_E_ARG, "%s: no HIP allocation holds this address", name);
    }
    const uintptr_t lo = reinterpret_cast<uintptr_t>(base);
    const uintptr_t at = reinterpret_cast<uintptr_t>(ptr);
    if (at < lo || at - lo > size || bytes > size - (at - lo))
        return fail(PF_E_ARG, "%s: %zu bytes run past the end of its allocation", name, bytes);
    return PF_OK;
}

// Makes ctx's device current for the scope of a device-pointer call and
// restores the caller's device afterwards (PyTorch tracks its own).
struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

// Grid launches are 1-D: n_curves * blocks-per-curve must fit one dimension.
int check_blocks(int64_t n_curves, int64_t per_curve) {
    if (per_curve > 0 && n_curves > (int64_t(INT32_MAX) / per_curve))
        return fail(PF_E_LIMIT, "too many curves x blocks for one launch");
    return PF_OK;
}

int64_t freq_blocks(int64_t K) {
    int64_t kb = 1;
    while (kb < K && kb < 64) kb <<= 1;
    return (K + kb - 1) / kb;
}

// LS grid over freqs (as f32 periods 1/f), device peaks, then Newton
// refinement of every peak inside its grid-neighbour bracket.
int ls_refine_on(pf_ctx* ctx, CurveSet& cs, const double* freqs, int64_t nf, int64_t n_peaks,
                 int64_t min_distance, double rtol, int max_iter, int64_t* idx, float* val,
                 double* f_out, double* p_out, int* status) {
    if (!freqs && nf > 0) return fail(PF_E_ARG, "freqs is NULL");
    if (nf < 0) return fail(PF_E_ARG, "n_freqs < 0");
    if (n_peaks < 0 || n_peaks > 64) return fail(PF_E_LIMIT, "n_peaks must be in [0, 64]");
    if (min_distance < 0) return fail(PF_E_ARG, "min_distance < 0");
    if (!(rtol > 0.0) || max_iter < 1) return fail(PF_E_ARG, "rtol must be > 0, max_iter >= 1");
    const int64_t C = cs.n_curves;
    const size_t pk = static_cast<size_t>(C) * n_peaks;
    if (pk == 0) return PF_OK;
    if (!idx || !val || !f_out || !p_out || !status) return fail(PF_E_ARG, "an output is NULL");
    std::vector<float> periods(static_cast<size_t>(nf));
    for (int64_t k = 0; k < nf; ++k) {
        if (!(freqs[k] > 0.0) || !std::isfinite(freqs[k]))
            return fail(PF_E_ARG, "freqs must be finite and > 0 (index %lld)",
                        static_cast<long long>(k));
        periods[k] = static_cast<float>(1.0 / freqs[k]);
    }
    if (nf < 3) {
        for (size_t i = 0; i < pk; ++i) {
            idx[i] = -1;
            val[i] = 0.0f;
            f_out[i] = 0.0;
            p_out[i] = 0.0;
            status[i] = 3;
        }
        return PF_OK;
    }
    PF_TRY(check_blocks(C, freq_blocks(n_peaks)));
    const float zero = 0.0f;
    Params p{PF_ALGO_LS, 0, 0, 0, 0, 0, 0.0f, 0.0f};
    PF_TRY(upload_grid(ctx, periods.data(), nf, &zero, 1));
    PF_TRY(run_grid(ctx, cs, p, nf, 1));
    hipStream_t s = ctx->stream;
    OutPlan plan;
    const size_t o_idx = plan.add(idx, pk * sizeof(int64_t));
    const size_t o_val = plan.add(val, pk * sizeof(float));
    const size_t o_w = plan.add(f_out, pk * sizeof(double));
    const size_t o_p = plan.add(p_out, pk * sizeof(double));
    const size_t o_st = plan.add(status, pk * sizeof(int));
    PF_TRY(ensure(ctx->outbuf, plan.total));
    int64_t* d_idx = out_dev<int64_t>(ctx, o_idx);
    double* d_w = out_dev<double>(ctx, o_w);
    double* d_p = out_dev<double>(ctx, o_p);
    int* d_st = out_dev<int>(ctx, o_st);
    pf::PeaksWorkspace ws;
    PF_TRY(peaks_workspace(ctx, C, nf, n_peaks, min_distance, &ws));
    PF_HIP(pf::launch_peaks(as<float>(cs.out), C, nf, n_peaks, min_distance, 1, d_idx,
                            out_dev<float>(ctx, o_val), s, &ws));
    PF_TRY(ensure(ctx->ls_freq, nf * sizeof(double)));
    PF_TRY(ensure(ctx->ls_lo, pk * sizeof(double)));
    PF_TRY(ensure(ctx->ls_hi, pk * sizeof(double)));
    PF_HIP(hipMemcpyAsync(ctx->ls_freq.p, freqs, nf * sizeof(double), hipMemcpyHostToDevice, s));
    PF_HIP(pf::launch_ls_bracket(d_idx, as<double>(ctx->ls_freq), nf, static_cast<int64_t>(pk),
                                 as<double>(ctx->ls_lo), as<double>(ctx->ls_hi), d_w, d_p, d_st,
                                 s));
    PF_HIP(pf::launch_ls_refine(as<int64_t>(cs.off), as<double>(cs.t64), cs.y.p, true, C, n_peaks,
                                as<double>(ctx->ls_lo), as<double>(ctx->ls_hi), d_w, d_p, d_st,
                                rtol, max_iter, s));
    PF_TRY(out_finish(ctx, plan));
    constexpr double kTwoPi = 6.283185307179586476925286766559;
    for (size_t i = 0; i < pk; ++i) f_out[i] /= kTwoPi;
    return PF_OK;
}

}  // namespace

extern "C" {

const char* pf_last_error(void) { return g_err.c_str(); }

int pf_device_count(int* count) {
    if (!count) return fail(PF_E_ARG, "count is NULL");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
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
    if (const char* g = std::getenv("PF_RMW_GROUP")) {
        const int v = std::atoi(g);
        ctx->rmw_group = (v >= 4) ? 4 : (v == 2 ? 2 : 1);
    }
    if (const char* g = std::getenv("PF_CE_GROUP")) {
        const int v = std::atoi(g);
        ctx->ce_group = (v >= 4) ? 4 : (v == 2 ? 2 : (v == 1 ? 1 : 0));
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
        cus > 0)
        ctx->n_cu = cus;
    ctx->arch = device_arch(device);
    {
        const char* env = std::getenv("PF_TUNE_TABLE");
        std::string path = env ? std::string(env) : default_tune_path(ctx->arch);
        if (!path.empty() && path != "none") {
            std::vector<pf_tune_entry> t;
            const int rc = parse_tune_file(path.c_str(), &t);
            if (rc != PF_OK) {
                delete ctx;
                return rc;
            }
            ctx->tune.swap(t);
            ctx->tune_source = path;
        }
    }
    // The allocation-free entry point's device array on this device (the
    // device is current); a failed lookup only disables that entry point.
    ctx->noalloc_rc = pf::noalloc_scratch(device, &ctx->noalloc);
    if (ctx->noalloc_rc != hipSuccess) (void)hipGetLastError();
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
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->scratch.release();
    DevBuf* bufs[] = {&ctx->periods, &ctx->pdts,   &ctx->outbuf, &ctx->ls_freq,
                      &ctx->ls_lo,   &ctx->ls_hi,  &ctx->peaks_ws, &ctx->hc_keep,
                      &ctx->hc_cnt,  &ctx->hc_src, &ctx->hc_off};
    for (DevBuf* b : bufs)
        if (b->p) (void)hipFree(b->p);
    for (pf_ctx::MfTable& m : ctx->mf_tables)
        if (m.buf.p) (void)hipFree(m.buf.p);
    if (ctx->host_stage) (void)hipHostFree(ctx->host_stage);
    for (int i = 0; i < 2; ++i) {
        if (ctx->in_stage[i]) (void)hipHostFree(ctx->in_stage[i]);
        if (ctx->in_ev[i]) (void)hipEventDestroy(ctx->in_ev[i]);
    }
    for (hipStream_t& st : ctx->side)
        if (st) (void)hipStreamDestroy(st);
    if (ctx->join) (void)hipEventDestroy(ctx->join);
    for (hipEvent_t ev : ctx->job_ev)
        if (ev) (void)hipEventDestroy(ev);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
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

int pf_mf_grid(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets, const float* t,
               const float* y, const float* err, const float* periods, int64_t n_periods,
               const float* period_dts, int64_t n_pdts, int num_bins, float* out) {
    Params p{PF_ALGO_MF, 0, 0, 1, 1, num_bins, 0.0f, 0.0f};
    return grid_host(ctx, p, n_curves, offsets, t, y, err, periods, n_periods, period_dts,
                     n_pdts, out);
}

int pf_vn_grid(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets, const float* t,
               const float* y, const float* periods, int64_t n_periods, const float* period_dts,
               int64_t n_pdts, int n_phase, int n_mag, int phase_ext, int mag_ext, int bandwidth,
               int margin, float* out) {
    Params p{PF_ALGO_VN, n_phase, n_mag, phase_ext, mag_ext, 0, 0.0f, 0.0f, PF_MAG_AS_IS,
             bandwidth, margin};
    return grid_host(ctx, p, n_curves, offsets, t, y, nullptr, periods, n_periods, period_dts,
                     n_pdts, out);
}

int pf_mf_features(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets, const float* t,
                   const float* y, const float* err, const float* periods, int num_bins,
                   float* out) {
    if (!ctx) return fail(PF_E_ARG, "ctx is NULL");
    Params p{PF_ALGO_MF, 0, 0, 1, 1, num_bins, 0.0f, 0.0f};
    PF_TRY(check_params(p, ctx->lds_max));
    if (n_curves > 0 && (!periods || !out)) return fail(PF_E_ARG, "periods/out is NULL");
    PF_HIP(hipSetDevice(ctx->device));
    PF_TRY(upload_curves(ctx, ctx->scratch, n_curves, offsets, t, y, err));
    CurveSet& cs = ctx->scratch;
    if (n_curves == 0) return finish(ctx);
    // One trial per curve: the per-curve periods are the trial grid (period_dt
    // 0), and the kernel folds curve c at periods[c].
    const float zero = 0.0f;
    PF_TRY(upload_grid(ctx, periods, n_curves, &zero, 1));
    PF_TRY(prepare(ctx, cs, p, false, ctx->stream));
    const size_t ob = static_cast<size_t>(n_curves) * pf::kMfFeatures * sizeof(float);
    PF_TRY(ensure(ctx->outbuf, ob));
    pf::GridArgs g;
    g.off = as<int64_t>(cs.off);
    g.t64 = as<double>(cs.t64);
    g.tt = as<double>(cs.t64);
    g.periods = as<float>(ctx->periods);
    g.pdts = as<float>(ctx->pdts);
    g.nd = 1;
    g.T = 1;
    g.tblocks = 1;
    g.n_blocks = static_cast<uint64_t>(n_curves);
    g.out = nullptr;
    g.has_pdt = 0;
    g.tmax = time_bound(cs, p);
    if (g.n_blocks > 0xffffffffull) return fail(PF_E_LIMIT, "too many curves for one launch");
    pf::MfArgs a{as<float>(cs.w), as<float>(cs.fpw_wy), as<float>(cs.mf_wyy), mags(cs, p),
                 mf_templates(ctx, num_bins), num_bins, as<float>(ctx->periods),
                 as<float>(ctx->outbuf)};
    PF_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    const hipError_t le = pf::launch_mf(g, a, pf::lds_mf(num_bins), 1, nullptr, ctx->stream);
    if (le != hipSuccess) return fail(PF_E_HIP, "MF features launch failed: %s", hipGetErrorString(le));
    PF_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    ctx->timed = true;
    PF_HIP(hipMemcpyAsync(out, ctx->outbuf.p, ob, hipMemcpyDeviceToHost, ctx->stream));
    return finish(ctx);
}

int pf_mhf_grid(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets, const float* t,
                const float* y, const float* err, const float* periods, int64_t n_periods,
                const float* period_dts, int64_t n_pdts, int max_harmonics, float* out) {
    Params p{PF_ALGO_MHF, 0, 0, 1, 1, 0, 0.0f, 0.0f};
    p.max_harmonics = max_harmonics;
    return grid_host(ctx, p, n_curves, offsets, t, y, err, periods, n_periods, period_dts,
                     n_pdts, out);
}

int pf_mhf_per_k(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets, const float* t,
                 const float* y, const float* err, const float* periods,
                 const float* period_dts, int max_harmonics, float* out) {
    if (!ctx) return fail(PF_E_ARG, "ctx is NULL");
    Params p{PF_ALGO_MHF, 0, 0, 1, 1, 0, 0.0f, 0.0f};
    p.max_harmonics = max_harmonics;
    PF_TRY(check_params(p, ctx->lds_max));
    if (n_curves > 0 && (!periods || !out)) return fail(PF_E_ARG, "periods/out is NULL");
    PF_HIP(hipSetDevice(ctx->device));
    PF_TRY(upload_curves(ctx, ctx->scratch, n_curves, offsets, t, y, err));
    CurveSet& cs = ctx->scratch;
    if (n_curves == 0) return finish(ctx);
    // One trial per curve: curve c is folded at (periods[c], period_dts[c]).
    std::vector<float> zeros;
    if (!period_dts) {
        zeros.assign(static_cast<size_t>(n_curves), 0.0f);
        period_dts = zeros.data();
    }
    PF_TRY(upload_grid(ctx, periods, n_curves, period_dts, n_curves));
    PF_TRY(prepare(ctx, cs, p, false, ctx->stream));
    const int K = mhf_max_k(max_harmonics);
    const size_t ob = static_cast<size_t>(n_curves) * (K + 2) * sizeof(float);
    PF_TRY(ensure(ctx->outbuf, ob));
    pf::GridArgs g;
    g.off = as<int64_t>(cs.off);
    g.t64 = as<double>(cs.t64);
    g.tt = as<double>(cs.t64);
    g.periods = as<float>(ctx->periods);
    g.pdts = as<float>(ctx->pdts);
    g.nd = 1;
    g.T = 1;
    g.tblocks = 1;
    g.n_blocks = static_cast<uint64_t>(n_curves);
    g.out = nullptr;
    g.has_pdt = ctx->has_pdt ? 1 : 0;
    g.tmax = nullptr;  // MHF folds phases for its sums, not bins
    if (g.n_blocks > 0xffffffffull) return fail(PF_E_LIMIT, "too many curves for one launch");
    pf::MhfArgs a{};
    PF_TRY(mhf_args(ctx, cs, K, 1, as<float>(ctx->outbuf), &a));
    PF_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    const hipError_t le = pf::launch_mhf(g, a, nullptr, ctx->stream);
    if (le != hipSuccess) return fail(PF_E_HIP, "MHF per-k launch failed: %s", hipGetErrorString(le));
    PF_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    ctx->timed = true;
    PF_HIP(hipMemcpyAsync(out, ctx->outbuf.p, ob, hipMemcpyDeviceToHost, ctx->stream));
    return finish(ctx);
}

int pf_fourier_features(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets, const float* t,
                        const float* y, const float* err, const float* periods, float* out) {
    if (!ctx) return fail(PF_E_ARG, "ctx is NULL");
    if (n_curves > 0 && (!periods || !out)) return fail(PF_E_ARG, "periods/out is NULL");
    PF_HIP(hipSetDevice(ctx->device));
    PF_TRY(upload_curves(ctx, ctx->scratch, n_curves, offsets, t, y, err));
    CurveSet& cs = ctx->scratch;
    if (n_curves == 0) return finish(ctx);
    const float zero = 0.0f;
    PF_TRY(upload_grid(ctx, periods, n_curves, &zero, 1));
    Params p{PF_ALGO_MHF, 0, 0, 1, 1, 0, 0.0f, 0.0f};
    p.max_harmonics = pf::kMhfMaxK;
    PF_TRY(prepare(ctx, cs, p, false, ctx->stream));  // the MHF point records and curve sums
    const size_t ob = static_cast<size_t>(n_curves) * pf::kFourierFeatures * sizeof(float);
    PF_TRY(ensure(ctx->outbuf, ob));
    PF_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    const hipError_t le =
        pf::launch_fourier(as<int64_t>(cs.off), as<pf::MhfPoint>(cs.mhf_pt),
                           as<double>(cs.mhf_sums), as<float>(ctx->periods),
                           as<float>(ctx->outbuf), n_curves, ctx->stream);
    if (le != hipSuccess)
        return fail(PF_E_HIP, "Fourier features launch failed: %s", hipGetErrorString(le));
    PF_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    ctx->timed = true;
    PF_HIP(hipMemcpyAsync(out, ctx->outbuf.p, ob, hipMemcpyDeviceToHost, ctx->stream));
    return finish(ctx);
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
    Params p = to_params(params);
    PF_TRY(check_params(p, ctx->lds_max));
    PF_TRY(check_grid(periods, n_periods, period_dts, n_pdts));
    if (n_peaks < 0 || n_peaks > 64) return fail(PF_E_LIMIT, "n_peaks must be in [0, 64]");
    PF_HIP(hipSetDevice(ctx->device));
    PF_TRY(upload_curves(ctx, ctx->scratch, n_curves, offsets, t, y, err));
    return peaks_on(ctx, ctx->scratch, p, periods, n_periods, period_dts, n_pdts, n_peaks,
                    min_distance, idx, val);
}

// pf_peaks without any device allocation: the derived per-point arrays and
// the periodogram live in the per-GPU array of pf_noalloc.hip, laid out per
// call; inputs and outputs are the caller's device memory.  Every region is
// written in full by the kernel that first reads it later in the same call,
// so nothing has to start at zero.  Calls that overlap on one GPU share the
// array and race.
int pf_peaks_noalloc(pf_ctx* ctx, void* stream, const pf_params* params, int64_t n_curves,
                     const int64_t* offsets, const float* t, const float* y, const float* err,
                     const float* periods, int64_t n_periods, const float* period_dts,
                     int64_t n_pdts, int64_t n_peaks, int64_t min_distance, int64_t* idx,
                     float* val) {
    if (!ctx || !params) return fail(PF_E_ARG, "ctx/params is NULL");
    const Params p = to_params(params);
    PF_TRY(check_params(p, ctx->lds_max));
    if (p.algo != PF_ALGO_CE && p.algo != PF_ALGO_AOV && p.algo != PF_ALGO_LS &&
        p.algo != PF_ALGO_FPW && p.algo != PF_ALGO_BLS)
        return fail(PF_E_ARG, "pf_peaks_noalloc evaluates CE, AOV, LS, FPW and BLS only");
    PF_TRY(check_grid(periods, n_periods, period_dts, n_pdts));
    if (n_peaks < 0 || n_peaks > 64) return fail(PF_E_LIMIT, "n_peaks must be in [0, 64]");
    if (min_distance < 0) return fail(PF_E_ARG, "min_distance < 0");
    int64_t npts = 0, mx = 0;
    PF_TRY(validate_offsets(n_curves, offsets, &npts, &mx));
    if (n_curves == 0 || n_peaks == 0) return PF_OK;

    // Scratch layout (256-byte aligned regions); sizes are checked against
    // the array before any product can overflow.
    constexpr uint64_t kMax = pf::kNoAllocScratchBytes;
    const uint64_t C = static_cast<uint64_t>(n_curves), N = static_cast<uint64_t>(npts);
    const uint64_t T = static_cast<uint64_t>(n_periods) * static_cast<uint64_t>(n_pdts);
    if (C > kMax / 8 || N > kMax / 8 || static_cast<uint64_t>(n_periods) > kMax / 4 ||
        static_cast<uint64_t>(n_pdts) > kMax / 4 || (T > 0 && C > kMax / 4 / T))
        return fail(PF_E_LIMIT, "pf_peaks_noalloc: the call exceeds the %llu-byte scratch array",
                    static_cast<unsigned long long>(kMax));
    bool has_pdt = false;
    for (int64_t i = 0; i < n_pdts; ++i)
        if (period_dts[i] != 0.0f) has_pdt = true;
    const bool prep = p.mag_prep != PF_MAG_AS_IS;
    uint64_t total = 0;
    auto region = [&](uint64_t bytes) {
        const uint64_t at = total;
        total += (std::max<uint64_t>(bytes, 16) + 255) & ~uint64_t(255);
        return at;
    };
    const uint64_t o_off = region((C + 1) * 8), o_per = region(n_periods * 4ull),
                   o_pdt = region(n_pdts * 4ull), o_t64 = region(N * 8),
                   o_tt = has_pdt ? region(N * 8) : 0, o_yp = prep ? region(N * 4) : 0,
                   o_tmax = region(C * 4);
    uint64_t o_a = 0, o_b = 0, o_c = 0;  // the statistic's own arrays
    switch (p.algo) {
        case PF_ALGO_CE:  // magnitude rows; packed columns (even n_mag)
            o_a = region(N * 4);
            if (p.n_mag % 2 == 0) o_b = region(N * 8);
            break;
        case PF_ALGO_AOV: o_a = region(N * 4); o_c = region(C * 4); break;  // y*y, mean
        case PF_ALGO_FPW: o_a = region(N * 4); o_b = region(N * 4); break;  // w, w*y
        case PF_ALGO_BLS:                                                  // w, yw, mean
            o_a = region(N * 4);
            o_b = region(N * 4);
            o_c = region(C * 4);
            break;
        default: break;
    }
    const uint64_t o_out = region(C * T * 4);
    if (total > kMax)
        return fail(PF_E_LIMIT,
                    "pf_peaks_noalloc: this call needs %llu bytes of scratch; the array holds %llu",
                    static_cast<unsigned long long>(total), static_cast<unsigned long long>(kMax));

    DeviceGuard dg(ctx->device);
    if (ctx->noalloc_rc != hipSuccess || !ctx->noalloc)
        return fail(PF_E_HIP, "the allocation-free scratch array is unavailable: %s",
                    hipGetErrorString(ctx->noalloc_rc));
    PF_TRY(check_dev_span(ctx, idx, C * n_peaks * sizeof(int64_t), "idx"));
    PF_TRY(check_dev_span(ctx, val, C * n_peaks * sizeof(float), "val"));
    if (N > 0) {
        PF_TRY(check_dev_span(ctx, t, N * 4, "t"));
        PF_TRY(check_dev_span(ctx, y, N * 4, "y"));
        if (err) PF_TRY(check_dev_span(ctx, err, N * 4, "err"));
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    char* base = static_cast<char*>(ctx->noalloc);

    // A curve set over the scratch regions and the caller's arrays; every
    // buffer is fixed, so nothing below can allocate.
    CurveSet cs;
    cs.fix_all();
    auto fixed = [](void* ptr, uint64_t bytes) {
        DevBuf b;
        b.p = ptr;
        b.cap = std::max<uint64_t>(bytes, 16);
        b.fixed = true;
        return b;
    };
    cs.n_curves = n_curves;
    cs.npts = npts;
    cs.max_len = mx;
    cs.has_err = err != nullptr;
    {
        double s1 = 0.0, s2 = 0.0;
        for (int64_t c = 0; c < n_curves; ++c) {
            const double len = static_cast<double>(offsets[c + 1] - offsets[c]);
            s1 += len;
            s2 += len * len;
        }
        cs.rep_len = s1 > 0.0 ? static_cast<int64_t>(s2 / s1) : 0;
    }
    cs.off = fixed(base + o_off, (C + 1) * 8);
    cs.t = fixed(const_cast<float*>(t), N * 4);
    cs.y = fixed(const_cast<float*>(y), N * 4);
    if (err) cs.e = fixed(const_cast<float*>(err), N * 4);
    cs.t64 = fixed(base + o_t64, N * 8);
    cs.tmax = fixed(base + o_tmax, C * 4);
    if (has_pdt) cs.tt = fixed(base + o_tt, N * 8);
    if (prep) cs.yp = fixed(base + o_yp, N * 4);
    switch (p.algo) {
        case PF_ALGO_CE:
            cs.ce_mrow = fixed(base + o_a, N * 4);
            if (p.n_mag % 2 == 0) cs.ce_mpk = fixed(base + o_b, N * 8);
            break;
        case PF_ALGO_AOV:
            cs.aov_yy = fixed(base + o_a, N * 4);
            cs.aov_avg = fixed(base + o_c, C * 4);
            break;
        case PF_ALGO_FPW:
            cs.w = fixed(base + o_a, N * 4);
            cs.fpw_wy = fixed(base + o_b, N * 4);
            break;
        case PF_ALGO_BLS:
            cs.w = fixed(base + o_a, N * 4);
            cs.bls_yw = fixed(base + o_b, N * 4);
            cs.bls_mean = fixed(base + o_c, C * 4);
            break;
        default: break;
    }
    cs.out = fixed(base + o_out, C * T * 4);
    cs.out_T = static_cast<int64_t>(T);

    // Host arrays in, then the same kernels as pf_peaks.
    PF_HIP(hipMemcpyAsync(cs.off.p, offsets, (C + 1) * 8, hipMemcpyHostToDevice, s));
    if (T > 0) {
        PF_HIP(hipMemcpyAsync(base + o_per, periods, n_periods * 4ull, hipMemcpyHostToDevice, s));
        PF_HIP(hipMemcpyAsync(base + o_pdt, period_dts, n_pdts * 4ull, hipMemcpyHostToDevice, s));
        const TrialGrid tg{reinterpret_cast<const float*>(base + o_per),
                           reinterpret_cast<const float*>(base + o_pdt), has_pdt};
        PF_TRY(prepare(ctx, cs, p, has_pdt, s));
        PF_TRY(launch_grid(ctx, cs, p, n_periods, n_pdts, tg, nullptr, false,
                           tune_choice(ctx, p.algo, cs.rep_len).group, s));
    }
    const int use_max = p.algo == PF_ALGO_CE ? 0 : 1;
    PF_HIP(pf::launch_peaks(as<float>(cs.out), n_curves, static_cast<int64_t>(T), n_peaks,
                            min_distance, use_max, idx, val, s));
    return PF_OK;
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
    const size_t pk = static_cast<size_t>(n_curves) * n_peaks;
    if (T == 0) {
        for (size_t i = 0; i < pk; ++i) {
            idx[i] = -1;
            val[i] = 0.0f;
        }
        return PF_OK;
    }
    PF_HIP(hipSetDevice(ctx->device));
    CurveSet& cs = ctx->scratch;
    cs.invalidate();
    const size_t db = static_cast<size_t>(n_curves) * T * sizeof(float);
    PF_TRY(ensure(cs.out, db));
    OutPlan plan;
    const size_t o_idx = plan.add(idx, pk * sizeof(int64_t));
    const size_t o_val = plan.add(val, pk * sizeof(float));
    PF_TRY(ensure(ctx->outbuf, plan.total));
    pf::PeaksWorkspace ws;
    PF_TRY(peaks_workspace(ctx, n_curves, T, n_peaks, min_distance, &ws));
    PF_HIP(hipMemcpyAsync(cs.out.p, data, db, hipMemcpyHostToDevice, ctx->stream));
    PF_HIP(pf::launch_peaks(as<float>(cs.out), n_curves, T, n_peaks, min_distance,
                            use_max ? 1 : 0, out_dev<int64_t>(ctx, o_idx),
                            out_dev<float>(ctx, o_val), ctx->stream, &ws));
    ctx->timed = false;
    return out_finish(ctx, plan);
}

int pf_curves_create(pf_ctx* ctx, pf_curves** out, int64_t n_curves, const int64_t* offsets,
                     const float* t, const float* y, const float* err) {
    if (!ctx || !out) return fail(PF_E_ARG, "ctx/out is NULL");
    *out = nullptr;
    PF_HIP(hipSetDevice(ctx->device));
    pf_curves* cv = new pf_curves();
    cv->ctx = ctx;
    int rc = upload_curves(ctx, cv->cs, n_curves, offsets, t, y, err);
    if (rc == PF_OK) rc = finish(ctx);
    if (rc != PF_OK) {
        cv->cs.release();
        delete cv;
        return rc;
    }
    *out = cv;
    return PF_OK;
}

int pf_curves_load(pf_ctx* ctx, pf_curves* cv, int64_t n_curves, const float* const* t,
                   const float* const* y, const float* const* err, const int64_t* lengths) {
    if (!ctx || !cv) return fail(PF_E_ARG, "ctx/cv is NULL");
    if (cv->ctx != ctx) return fail(PF_E_ARG, "the curve set belongs to another context");
    PF_HIP(hipSetDevice(ctx->device));
    PF_TRY(upload_lists(ctx, cv->cs, n_curves, t, y, err, lengths));
    return finish(ctx);
}

int pf_curves_remove_high_cadence(pf_ctx* ctx, pf_curves* cv, float cadence_days,
                                  int64_t* kept_offsets, float* t, float* y, float* err) {
    if (!ctx || !cv) return fail(PF_E_ARG, "ctx/cv is NULL");
    if (cv->ctx != ctx) return fail(PF_E_ARG, "the curve set belongs to another context");
    CurveSet& cs = cv->cs;
    if (cs.max_len > static_cast<int64_t>(UINT32_MAX))
        return fail(PF_E_LIMIT, "a curve of %lld points exceeds the filter's 2^32 - 1",
                    static_cast<long long>(cs.max_len));
    PF_HIP(hipSetDevice(ctx->device));
    ctx->timed = false;
    const int64_t nc = cs.n_curves, np = cs.npts;
    const int planes = cs.has_err ? 3 : 2;
    PF_TRY(ensure(ctx->hc_keep, static_cast<size_t>(np) * sizeof(uint32_t)));
    PF_TRY(ensure(ctx->hc_cnt, static_cast<size_t>(nc) * sizeof(int64_t)));
    PF_TRY(ensure(ctx->hc_src, static_cast<size_t>(planes) * np * sizeof(float)));
    PF_TRY(ensure(ctx->hc_off, static_cast<size_t>(nc + 1) * sizeof(int64_t)));
    hipStream_t s = ctx->stream;
    PF_HIP(pf::launch_cadence_keep(as<int64_t>(cs.off), as<float>(cs.t), cadence_days,
                                   as<uint32_t>(ctx->hc_keep), as<int64_t>(ctx->hc_cnt), nc, s));
    std::vector<int64_t> cnt(static_cast<size_t>(nc));
    if (nc) PF_HIP(hipMemcpyAsync(cnt.data(), ctx->hc_cnt.p, nc * sizeof(int64_t),
                                  hipMemcpyDeviceToHost, s));
    PF_HIP(hipStreamSynchronize(s));
    std::vector<int64_t> noff(static_cast<size_t>(nc) + 1, 0);
    int64_t mx = 0;
    for (int64_t c = 0; c < nc; ++c) {
        noff[c + 1] = noff[c] + cnt[c];
        if (cnt[c] > mx) mx = cnt[c];
    }
    const int64_t kept = noff[nc];
    // The set's planes are copied aside, then gathered back compacted.
    float* src = as<float>(ctx->hc_src);
    const size_t nb = static_cast<size_t>(np) * sizeof(float);
    if (np) {
        PF_HIP(hipMemcpyAsync(src, cs.t.p, nb, hipMemcpyDeviceToDevice, s));
        PF_HIP(hipMemcpyAsync(src + np, cs.y.p, nb, hipMemcpyDeviceToDevice, s));
        if (cs.has_err)
            PF_HIP(hipMemcpyAsync(src + 2 * np, cs.e.p, nb, hipMemcpyDeviceToDevice, s));
    }
    PF_HIP(hipMemcpyAsync(ctx->hc_off.p, noff.data(), (nc + 1) * sizeof(int64_t),
                          hipMemcpyHostToDevice, s));
    // From the compaction on, the planes no longer match the old offsets: a
    // failure there leaves an empty, consistent set.
    auto compact = [&]() -> int {
        PF_HIP(pf::launch_cadence_gather(as<int64_t>(cs.off), as<int64_t>(ctx->hc_off),
                                         as<uint32_t>(ctx->hc_keep), src, np, as<float>(cs.t),
                                         as<float>(cs.y), cs.has_err ? as<float>(cs.e) : nullptr,
                                         nc, s));
        PF_HIP(hipMemcpyAsync(cs.off.p, ctx->hc_off.p, (nc + 1) * sizeof(int64_t),
                              hipMemcpyDeviceToDevice, s));
        return PF_OK;
    };
    if (const int rc = compact(); rc != PF_OK) {
        const int64_t zero = 0;
        set_layout(cs, 0, &zero, 0, 0, false);
        return rc;
    }
    set_layout(cs, nc, noff.data(), kept, mx, cs.has_err);  // drops the derived arrays
    if (kept_offsets) std::copy(noff.begin(), noff.end(), kept_offsets);
    const size_t kb = static_cast<size_t>(kept) * sizeof(float);
    if (kb) {
        if (t) PF_HIP(hipMemcpyAsync(t, cs.t.p, kb, hipMemcpyDeviceToHost, s));
        if (y) PF_HIP(hipMemcpyAsync(y, cs.y.p, kb, hipMemcpyDeviceToHost, s));
        if (err && cs.has_err) PF_HIP(hipMemcpyAsync(err, cs.e.p, kb, hipMemcpyDeviceToHost, s));
    }
    PF_TRY(finish(ctx));
    if (err && !cs.has_err) std::fill(err, err + kept, 1.0f);  // a set without errors: ones
    return PF_OK;
}

void pf_curves_destroy(pf_curves* cv) {
    if (!cv) return;
    (void)hipSetDevice(cv->ctx->device);
    (void)hipStreamSynchronize(cv->ctx->stream);
    cv->cs.release();
    delete cv;
}

int pf_curves_grid(pf_ctx* ctx, pf_curves* cv, const pf_params* params, const float* periods,
                   int64_t n_periods, const float* period_dts, int64_t n_pdts, float* out) {
    if (!ctx || !cv || !params) return fail(PF_E_ARG, "NULL argument");
    if (cv->ctx != ctx) return fail(PF_E_ARG, "curve set belongs to another context");
    PF_HIP(hipSetDevice(ctx->device));
    return grid_on(ctx, cv->cs, to_params(params), periods, n_periods, period_dts, n_pdts, out);
}

int pf_curves_result(pf_curves* cv, const float** out_d, int64_t* n_curves, int64_t* T) {
    if (!cv || !out_d || !n_curves || !T) return fail(PF_E_ARG, "NULL argument");
    *out_d = as<float>(cv->cs.out);
    *n_curves = cv->cs.n_curves;
    *T = cv->cs.out_T;
    return PF_OK;
}

int pf_curves_peaks(pf_ctx* ctx, pf_curves* cv, const pf_params* params, const float* periods,
                    int64_t n_periods, const float* period_dts, int64_t n_pdts,
                    int64_t n_peaks, int64_t min_distance, int64_t* idx, float* val) {
    if (!ctx || !cv || !params) return fail(PF_E_ARG, "NULL argument");
    if (cv->ctx != ctx) return fail(PF_E_ARG, "curve set belongs to another context");
    PF_HIP(hipSetDevice(ctx->device));
    return peaks_on(ctx, cv->cs, to_params(params), periods, n_periods, period_dts, n_pdts,
                    n_peaks, min_distance, idx, val);
}

int pf_stats(pf_ctx* ctx, const pf_params* params, int64_t n_curves, const int64_t* offsets,
             const float* t, const float* y, const float* err, const float* periods,
             int64_t n_periods, const float* period_dts, int64_t n_pdts, int n_stats,
             float* summary, int64_t* idx, float* val, int* fallback) {
    if (!ctx || !params) return fail(PF_E_ARG, "ctx/params is NULL");
    Params p = to_params(params);
    PF_TRY(check_params(p, ctx->lds_max));
    PF_TRY(check_grid(periods, n_periods, period_dts, n_pdts));
    PF_HIP(hipSetDevice(ctx->device));
    PF_TRY(upload_curves(ctx, ctx->scratch, n_curves, offsets, t, y, err));
    return stats_on(ctx, ctx->scratch, p, periods, n_periods, period_dts, n_pdts, n_stats,
                    summary, idx, val, fallback);
}

int pf_curves_stats(pf_ctx* ctx, pf_curves* cv, const pf_params* params, const float* periods,
                    int64_t n_periods, const float* period_dts, int64_t n_pdts, int n_stats,
                    float* summary, int64_t* idx, float* val, int* fallback) {
    if (!ctx || !cv || !params) return fail(PF_E_ARG, "NULL argument");
    if (cv->ctx != ctx) return fail(PF_E_ARG, "curve set belongs to another context");
    PF_HIP(hipSetDevice(ctx->device));
    return stats_on(ctx, cv->cs, to_params(params), periods, n_periods, period_dts, n_pdts,
                    n_stats, summary, idx, val, fallback);
}

int pf_curves_stats_many(pf_ctx* ctx, int64_t n_jobs, pf_curves* const* sets,
                         const pf_params* params, const float* periods, int64_t n_periods,
                         const float* period_dts, int64_t n_pdts, int n_stats,
                         float* const* summary, int64_t* const* idx, float* const* val,
                         int* const* fallback, float* kernel_ms) {
    if (!ctx) return fail(PF_E_ARG, "ctx is NULL");
    if (n_jobs < 0) return fail(PF_E_ARG, "n_jobs < 0");
    if (n_jobs == 0) return PF_OK;
    if (!sets || !params || !summary || !idx || !val || !fallback)
        return fail(PF_E_ARG, "NULL job array");
    PF_TRY(check_grid(periods, n_periods, period_dts, n_pdts));
    const int64_t T = n_periods * n_pdts;
    if (T < 1 || T > pf::kStatsMaxT)
        return fail(PF_E_LIMIT, "device stats need 1 <= n_periods*n_pdts <= %lld",
                    static_cast<long long>(pf::kStatsMaxT));
    if (n_stats < 1 || n_stats > T) return fail(PF_E_ARG, "n_stats must be in [1, T]");
    if (pf::stats_lds_bytes(T) + pf::stats_lds_static_bytes() > ctx->lds_max)
        return fail(PF_E_LIMIT, "stats exceed LDS");
    for (int64_t j = 0; j < n_jobs; ++j) {
        if (!sets[j]) return fail(PF_E_ARG, "job %lld: curve set is NULL", static_cast<long long>(j));
        if (sets[j]->ctx != ctx)
            return fail(PF_E_ARG, "job %lld: curve set belongs to another context",
                        static_cast<long long>(j));
        for (int64_t i = 0; i < j; ++i)
            if (sets[i] == sets[j])
                return fail(PF_E_ARG, "jobs %lld and %lld share a curve set",
                            static_cast<long long>(i), static_cast<long long>(j));
        PF_TRY(check_params(to_params(&params[j]), ctx->lds_max));
        if (sets[j]->cs.n_curves > 0 &&
            (!summary[j] || !idx[j] || !val[j] || !fallback[j]))
            return fail(PF_E_ARG, "job %lld: NULL output", static_cast<long long>(j));
    }
    PF_HIP(hipSetDevice(ctx->device));
    PF_TRY(upload_grid(ctx, periods, n_periods, period_dts, n_pdts));
    // Streams and events, once per context.
    for (hipStream_t& st : ctx->side)
        if (!st) PF_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    if (!ctx->join) PF_HIP(hipEventCreateWithFlags(&ctx->join, hipEventDisableTiming));
    const bool timing = kernel_ms != nullptr;  // per-job events only when asked for
    while (timing && static_cast<int64_t>(ctx->job_ev.size()) < 2 * n_jobs) {
        hipEvent_t ev = nullptr;
        PF_HIP(hipEventCreate(&ev));
        ctx->job_ev.push_back(ev);
    }
    // Derived arrays and result buffers, on the context stream (cached
    // after the first call).
    OutPlan plan;
    std::vector<size_t> o_idx(n_jobs), o_val(n_jobs), o_sum(n_jobs), o_flg(n_jobs);
    for (int64_t j = 0; j < n_jobs; ++j) {
        CurveSet& cs = sets[j]->cs;
        if (cs.n_curves == 0) continue;
        const Params p = to_params(&params[j]);
        PF_TRY(prepare(ctx, cs, p, ctx->has_pdt, ctx->stream));
        PF_TRY(prepare_equal_weights(ctx, cs, p, T, false, n_jobs > 1, ctx->stream));
        cs.out_T = 0;
        PF_TRY(ensure(cs.out, static_cast<size_t>(cs.n_curves) * T * sizeof(float)));
        cs.out_T = T;
        const size_t C = static_cast<size_t>(cs.n_curves);
        o_idx[j] = plan.add(idx[j], C * n_stats * sizeof(int64_t));
        o_val[j] = plan.add(val[j], C * n_stats * sizeof(float));
        o_sum[j] = plan.add(summary[j], C * 4 * sizeof(float));
        o_flg[j] = plan.add(fallback[j], C * sizeof(int));
    }
    PF_TRY(ensure(ctx->outbuf, plan.total));
    PF_HIP(hipEventRecord(ctx->join, ctx->stream));
    const int K = pf_ctx::kSideStreams;
    for (int k = 0; k < K; ++k) PF_HIP(hipStreamWaitEvent(ctx->side[k], ctx->join, 0));
    // Streams: the context stream and the side streams (four hardware
    // queues).  Jobs go longest first, by points x trials, to the
    // least-loaded stream; jobs of equal size keep the caller's order.  (A
    // per-statistic cost weight was measured 5% slower on the 15-job batch:
    // it started LDS-heavy launches side by side.)
    hipStream_t streams[K + 1] = {ctx->stream};
    for (int k = 0; k < K; ++k) streams[k + 1] = ctx->side[k];
    std::vector<double> cost(n_jobs);
    std::vector<int64_t> order(n_jobs);
    for (int64_t j = 0; j < n_jobs; ++j) {
        order[j] = j;
        cost[j] = static_cast<double>(sets[j]->cs.npts) * static_cast<double>(T);
    }
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t a, int64_t b) { return cost[a] > cost[b]; });
    // Until every stream has been joined, an early error return drains the
    // side streams first, so no kernel of this call is still reading buffers
    // that a later call may reuse.
    struct Drain {
        pf_ctx* ctx;
        bool armed = true;
        ~Drain() {
            if (!armed) return;
            for (hipStream_t st : ctx->side)
                if (st) (void)hipStreamSynchronize(st);
        }
    } drain{ctx};
    double load[K + 1] = {};
    for (const int64_t j : order) {
        CurveSet& cs = sets[j]->cs;
        int k = 0;
        for (int i = 1; i <= K; ++i)
            if (load[i] < load[k]) k = i;
        load[k] += cost[j];
        hipStream_t s = streams[k];
        if (timing) PF_HIP(hipEventRecord(ctx->job_ev[2 * j], s));
        if (cs.n_curves > 0) {
            const Params p = to_params(&params[j]);
            PF_TRY(launch_grid(ctx, cs, p, n_periods, n_pdts, ctx_grid(ctx), nullptr, false,
                               tune_choice(ctx, p.algo, cs.rep_len).group, s, n_jobs > 1));
            const int use_max = p.algo == PF_ALGO_CE ? 0 : 1;
            PF_HIP(pf::launch_stats(as<float>(cs.out), cs.n_curves, T, n_stats, use_max,
                                    out_dev<float>(ctx, o_sum[j]), out_dev<int64_t>(ctx, o_idx[j]),
                                    out_dev<float>(ctx, o_val[j]), out_dev<int>(ctx, o_flg[j]), s));
        }
        if (timing) PF_HIP(hipEventRecord(ctx->job_ev[2 * j + 1], s));
    }
    for (int k = 0; k < K; ++k) {
        PF_HIP(hipEventRecord(ctx->join, ctx->side[k]));
        PF_HIP(hipStreamWaitEvent(ctx->stream, ctx->join, 0));
    }
    drain.armed = false;  // the context stream now orders everything after this call
    ctx->timed = false;
    PF_TRY(out_finish(ctx, plan));
    if (kernel_ms)
        for (int64_t j = 0; j < n_jobs; ++j) {
            float ms = 0.0f;
            PF_HIP(hipEventElapsedTime(&ms, ctx->job_ev[2 * j], ctx->job_ev[2 * j + 1]));
            kernel_ms[j] = ms;
        }
    return PF_OK;
}

int pf_curves_peaks_ragged(pf_ctx* ctx, pf_curves* cv, const pf_params* params,
                           const float* periods, int64_t n_periods, const float* period_dts,
                           int64_t n_pdts, int64_t n_peaks, int64_t min_distance, int64_t* idx,
                           float* val) {
    if (!ctx || !cv || !params) return fail(PF_E_ARG, "NULL argument");
    if (cv->ctx != ctx) return fail(PF_E_ARG, "curve set belongs to another context");
    PF_HIP(hipSetDevice(ctx->device));
    return ragged_peaks_on(ctx, cv->cs, to_params(params), periods, n_periods, period_dts,
                           n_pdts, n_peaks, min_distance, idx, val);
}

int pf_ls_jet(pf_ctx* ctx, void* stream, int64_t n_curves, const int64_t* offsets,
              const double* t, const double* y, const double* omega, int64_t K, double* p,
              double* dp, double* d2p, double* coef) {
    if (!ctx) return fail(PF_E_ARG, "ctx is NULL");
    if (n_curves < 0 || K < 0) return fail(PF_E_ARG, "negative size");
    if (n_curves == 0 || K == 0) return PF_OK;
    PF_TRY(check_blocks(n_curves, freq_blocks(K)));
    DeviceGuard dg(ctx->device);
    PF_TRY(check_dev(ctx, offsets, "offsets"));
    PF_TRY(check_dev(ctx, t, "t"));
    PF_TRY(check_dev(ctx, y, "y"));
    PF_TRY(check_dev(ctx, omega, "omega"));
    PF_TRY(check_dev(ctx, p, "p"));
    PF_TRY(check_dev_opt(ctx, dp, "dp"));
    PF_TRY(check_dev_opt(ctx, d2p, "d2p"));
    PF_TRY(check_dev_opt(ctx, coef, "coef"));
    PF_HIP(pf::launch_ls_jet(offsets, t, y, false, omega, n_curves, K, p, dp, d2p, coef,
                             static_cast<hipStream_t>(stream)));
    return PF_OK;
}

int pf_ls_backward(pf_ctx* ctx, void* stream, int64_t n_curves, const int64_t* offsets,
                   int64_t max_len, const double* t, const double* omega, int64_t K,
                   const double* grad_p, const double* coef, const double* dp, double* grad_y,
                   double* grad_omega) {
    if (!ctx) return fail(PF_E_ARG, "ctx is NULL");
    if (n_curves < 0 || K < 0 || max_len < 0) return fail(PF_E_ARG, "negative size");
    if (n_curves == 0) return PF_OK;
    const int64_t pblocks = max_len > 0 ? (max_len + 255) / 256 : 1;
    PF_TRY(check_blocks(n_curves, pblocks));
    DeviceGuard dg(ctx->device);
    PF_TRY(check_dev(ctx, offsets, "offsets"));
    PF_TRY(check_dev(ctx, t, "t"));
    PF_TRY(check_dev(ctx, grad_y, "grad_y"));
    if (K > 0) {
        PF_TRY(check_dev(ctx, omega, "omega"));
        PF_TRY(check_dev(ctx, grad_p, "grad_p"));
        PF_TRY(check_dev(ctx, coef, "coef"));
    }
    if (grad_omega && K > 0) {
        PF_TRY(check_dev(ctx, grad_omega, "grad_omega"));
        PF_TRY(check_dev(ctx, dp, "dp"));
    }
    PF_HIP(pf::launch_ls_backward(offsets, t, omega, n_curves, max_len, K, grad_p, coef, dp,
                                  grad_y, K > 0 ? grad_omega : nullptr,
                                  static_cast<hipStream_t>(stream)));
    return PF_OK;
}

int pf_ls_refine(pf_ctx* ctx, void* stream, int64_t n_curves, const int64_t* offsets,
                 const double* t, const double* y, int64_t K, const double* lo, const double* hi,
                 double* omega, double* p, int* status, double rtol, int max_iter) {
    if (!ctx) return fail(PF_E_ARG, "ctx is NULL");
    if (n_curves < 0 || K < 0) return fail(PF_E_ARG, "negative size");
    if (!(rtol > 0.0) || max_iter < 1) return fail(PF_E_ARG, "rtol must be > 0, max_iter >= 1");
    if (n_curves == 0 || K == 0) return PF_OK;
    PF_TRY(check_blocks(n_curves, freq_blocks(K)));
    DeviceGuard dg(ctx->device);
    PF_TRY(check_dev(ctx, offsets, "offsets"));
    PF_TRY(check_dev(ctx, t, "t"));
    PF_TRY(check_dev(ctx, y, "y"));
    PF_TRY(check_dev(ctx, lo, "lo"));
    PF_TRY(check_dev(ctx, hi, "hi"));
    PF_TRY(check_dev(ctx, omega, "omega"));
    PF_TRY(check_dev(ctx, p, "p"));
    PF_TRY(check_dev(ctx, status, "status"));
    PF_HIP(pf::launch_ls_refine(offsets, t, y, false, n_curves, K, lo, hi, omega, p, status, rtol,
                                max_iter, static_cast<hipStream_t>(stream)));
    return PF_OK;
}

int pf_ls_peaks_refined(pf_ctx* ctx, int64_t n_curves, const int64_t* offsets, const float* t,
                        const float* y, const double* freqs, int64_t n_freqs, int64_t n_peaks,
                        int64_t min_distance, double rtol, int max_iter, int64_t* idx,
                        float* val, double* freq_out, double* power_out, int* status) {
    if (!ctx) return fail(PF_E_ARG, "ctx is NULL");
    PF_HIP(hipSetDevice(ctx->device));
    PF_TRY(upload_curves(ctx, ctx->scratch, n_curves, offsets, t, y, nullptr));
    return ls_refine_on(ctx, ctx->scratch, freqs, n_freqs, n_peaks, min_distance, rtol, max_iter,
                        idx, val, freq_out, power_out, status);
}

int pf_curves_ls_peaks_refined(pf_ctx* ctx, pf_curves* cv, const double* freqs, int64_t n_freqs,
                               int64_t n_peaks, int64_t min_distance, double rtol, int max_iter,
                               int64_t* idx, float* val, double* freq_out, double* power_out,
                               int* status) {
    if (!ctx || !cv) return fail(PF_E_ARG, "ctx or curve set is NULL");
    if (cv->ctx != ctx) return fail(PF_E_ARG, "curve set belongs to another context");
    PF_HIP(hipSetDevice(ctx->device));
    return ls_refine_on(ctx, cv->cs, freqs, n_freqs, n_peaks, min_distance, rtol, max_iter, idx,
                        val, freq_out, power_out, status);
}

int pf_ctx_set_tuning(pf_ctx* ctx, const pf_tune_entry* entries, int64_t n) {
    if (!ctx) return fail(PF_E_ARG, "ctx is NULL");
    PF_TRY(validate_tune(entries, n));
    ctx->tune.assign(entries, entries + n);
    ctx->tune_source = n ? "set" : "";
    return PF_OK;
}

int pf_ctx_get_tuning(pf_ctx* ctx, pf_tune_entry* entries, int64_t cap, int64_t* n) {
    if (!ctx || !n) return fail(PF_E_ARG, "ctx or n is NULL");
    const int64_t have = static_cast<int64_t>(ctx->tune.size());
    *n = have;
    if (cap > 0 && !entries) return fail(PF_E_ARG, "entries is NULL");
    for (int64_t i = 0; i < std::min(cap, have); ++i) entries[i] = ctx->tune[i];
    return PF_OK;
}

int pf_ctx_load_tuning(pf_ctx* ctx, const char* path) {
    if (!ctx || !path) return fail(PF_E_ARG, "ctx or path is NULL");
    std::vector<pf_tune_entry> t;
    PF_TRY(parse_tune_file(path, &t));
    ctx->tune.swap(t);
    ctx->tune_source = path;
    return PF_OK;
}

const char* pf_ctx_tuning_source(pf_ctx* ctx) {
    return ctx ? ctx->tune_source.c_str() : "";
}

const char* pf_ctx_arch(pf_ctx* ctx) { return ctx ? ctx->arch.c_str() : ""; }

int pf_last_kernel_ms(pf_ctx* ctx, float* ms) {
    if (!ctx || !ms) return fail(PF_E_ARG, "NULL argument");
    *ms = ctx->last_ms;
    return PF_OK;
}

}  // extern "C"
