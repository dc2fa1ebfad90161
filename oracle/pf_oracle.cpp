// CPU oracle for the periodogram hot path.
//
// TEST INFRASTRUCTURE ONLY. This file restates periodfind's Rust CPU backend
// (rust/src/{fold,ce,aov,ls,fpw,bls,peaks}.rs) in C++ so that the HIP path can
// be checked against it.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and there only as the checker or as the timed
// CPU baseline; the product path never calls into it.
//
// Arithmetic follows the Rust source operation by operation:
//   * every per-trial loop over points runs serially in point order with f32
//     accumulators (Rayon only parallelises over independent trials/curves);
//   * the fold is promoted to f64 in Rust's operation order;
//   * built with -ffp-contract=off so no multiply-add is fused (rustc does not
//     contract either);
//   * f32 sin/cos/ln/atan2 call the platform libm, as Rust std does on Linux.
// Parallelism: OpenMP over (curve, trial) pairs, the same decomposition as the
// Rust/Rayon path (rust/src/lib.rs:64-83 over curves, e.g. ce.rs:35-37 over
// trials), so it doubles as the CPU baseline.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <limits>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// Rust `f32 as usize`: saturating, NaN -> 0.
inline uint64_t f32_to_usize(float x) {
    if (!(x > 0.0f)) return 0;
    if (x >= 18446744073709551616.0f) return UINT64_MAX;
    return static_cast<uint64_t>(x);
}

// rust/src/fold.rs:13-15
inline float pdt_correction(float period, float period_dt) {
    return (period_dt / period) / 2.0f;
}

// rust/src/fold.rs:22-26: ((t - (k*t)*t) / P) in f64, |r - floor(r)| -> f32.
inline float fold_time(float t, float period, float pdt_corr) {
    double td = static_cast<double>(t);
    double t_corr = td - static_cast<double>(pdt_corr) * td * td;
    double ratio = t_corr / static_cast<double>(period);
    return static_cast<float>(std::fabs(ratio - std::floor(ratio)));
}

struct Grid {
    const float* periods;
    int64_t n_periods;
    const float* pdts;
    int64_t n_pdts;
};

// rust/src/ce.rs:13-91 for one (period, pdt) trial.
float ce_trial(const float* t, const float* m, int64_t n, float period, float pdt,
               int n_phase, int n_mag, int p_ext, int m_ext, uint32_t* hist) {
    const float phase_bin_size = 1.0f / static_cast<float>(n_phase);
    const float mag_bin_size = 1.0f / static_cast<float>(n_mag);
    const float norm = static_cast<float>(static_cast<uint64_t>(n) *
                                          static_cast<uint64_t>(p_ext) *
                                          static_cast<uint64_t>(m_ext));
    const float k = pdt_correction(period, pdt);
    std::memset(hist, 0, sizeof(uint32_t) * n_phase * n_mag);
    for (int64_t i = 0; i < n; ++i) {
        float ph = fold_time(t[i], period, k);
        uint64_t pb0 = f32_to_usize(ph / phase_bin_size);
        uint64_t mb0 = f32_to_usize(m[i] / mag_bin_size);
        for (int po = 0; po < p_ext; ++po) {
            for (int mo = 0; mo < m_ext; ++mo) {
                uint64_t pb = (pb0 + po) % static_cast<uint64_t>(n_phase);
                uint64_t mb = (mb0 + mo) % static_cast<uint64_t>(n_mag);
                hist[pb * n_mag + mb] += 1;
            }
        }
    }
    float ce = 0.0f;
    for (int j = 0; j < n_phase; ++j) {
        float pj = 0.0f;
        for (int i = 0; i < n_mag; ++i) pj += static_cast<float>(hist[j * n_mag + i]) / norm;
        for (int i = 0; i < n_mag; ++i) {
            float pij = static_cast<float>(hist[j * n_mag + i]) / norm;
            if (pij != 0.0f) {
                float l = logf(pj / pij);
                float term = pij * l;
                ce += term;
            }
        }
    }
    return ce;
}

// rust/src/aov.rs:12-89 for one trial; `avg` is the serial f32 mean (aov.rs:27).
float aov_trial(const float* t, const float* m, int64_t n, float avg, float period, float pdt,
                int n_bins, int ext, uint32_t* cnt, float* sum, float* sq) {
    const float bin_size = 1.0f / static_cast<float>(n_bins);
    const float k = pdt_correction(period, pdt);
    for (int b = 0; b < n_bins; ++b) { cnt[b] = 0; sum[b] = 0.0f; sq[b] = 0.0f; }
    for (int64_t i = 0; i < n; ++i) {
        float ph = fold_time(t[i], period, k);
        uint64_t b0 = f32_to_usize(ph / bin_size);
        float y = m[i];
        float yy = y * y;
        for (int o = 0; o < ext; ++o) {
            uint64_t b = (b0 + o) % static_cast<uint64_t>(n_bins);
            cnt[b] += 1;
            sum[b] += y;
            sq[b] += yy;
        }
    }
    float s1 = 0.0f, s2 = 0.0f;
    for (int b = 0; b < n_bins; ++b) {
        float c = static_cast<float>(cnt[b]);
        if (c != 0.0f) {
            float aux = sum[b] / c;
            float res = aux - avg;
            float a = c * res;
            a = a * res;
            s1 += a;
            float q = c * aux;
            q = q * aux;
            s2 += sq[b] - q;
        }
    }
    if (s2 == 0.0f) return 0.0f;
    float lhs = (static_cast<float>(n) - static_cast<float>(n_bins)) /
                (static_cast<float>(n_bins) - 1.0f);
    return lhs * (s1 / s2);
}

// rust/src/ls.rs:15-79 for one trial.
float ls_trial(const float* t, const float* y, int64_t n, float period, float pdt) {
    const float two_pi = 2.0f * static_cast<float>(M_PI);
    const float k = pdt_correction(period, pdt);
    float mc = 0.0f, ms = 0.0f, cc = 0.0f, cs = 0.0f;
    for (int64_t i = 0; i < n; ++i) {
        float ph = fold_time(t[i], period, k);
        float ang = two_pi * ph;
        float s = sinf(ang);
        float c = cosf(ang);
        float a;
        a = y[i] * c; mc += a;
        a = y[i] * s; ms += a;
        a = c * c;    cc += a;
        a = c * s;    cs += a;
    }
    float ss = static_cast<float>(n) - cc;
    float tau = 0.5f * atan2f(2.0f * cs, cc - ss);
    float st = sinf(tau), ct = cosf(tau);
    float v = ct * mc + st * ms;
    float num_l = v * v;
    v = ct * ms - st * mc;
    float num_r = v * v;
    float den_l = ct * ct * cc + 2.0f * ct * st * cs + st * st * ss;
    float den_r = ct * ct * ss - 2.0f * ct * st * cs + st * st * cc;
    if (den_l == 0.0f || den_r == 0.0f) return 0.0f;
    return 0.5f * (num_l / den_l + num_r / den_r);
}

// rust/src/fpw.rs:13-67 for one trial; w = 1/(e*e), wy = w*y precomputed (fpw.rs:26-27).
float fpw_trial(const float* t, const float* w, const float* wy, int64_t n, float period,
                float pdt, int n_bins, float* vw, float* vy) {
    const float k = pdt_correction(period, pdt);
    const float nbf = static_cast<float>(n_bins);
    for (int b = 0; b < n_bins; ++b) { vw[b] = 0.0f; vy[b] = 0.0f; }
    for (int64_t i = 0; i < n; ++i) {
        float ph = fold_time(t[i], period, k);
        uint64_t b = std::min<uint64_t>(f32_to_usize(ph * nbf), static_cast<uint64_t>(n_bins - 1));
        vw[b] += w[i];
        vy[b] += wy[i];
    }
    float s = 0.0f;
    for (int b = 0; b < n_bins; ++b) {
        if (vw[b] > 0.0f) {
            float num = vy[b] * vy[b];
            float den = 2.0f * vw[b];
            s += num / den;
        }
    }
    return s;
}

// rust/src/bls.rs:20-134 for one trial; w = 1/e^2, yw = w*(y - mean_w) precomputed (bls.rs:34-47).
float bls_trial(const float* t, const float* w, const float* yw, int64_t n, float period,
                float pdt, int n_bins, int64_t nb_min, int64_t nb_max, float* wb, float* yb,
                float* wp, float* yp) {
    const float k = pdt_correction(period, pdt);
    const float nbf = static_cast<float>(n_bins);
    for (int b = 0; b < n_bins; ++b) { wb[b] = 0.0f; yb[b] = 0.0f; }
    for (int64_t i = 0; i < n; ++i) {
        float ph = fold_time(t[i], period, k);
        uint64_t b = std::min<uint64_t>(f32_to_usize(ph * nbf), static_cast<uint64_t>(n_bins - 1));
        wb[b] += w[i];
        yb[b] += yw[i];
    }
    wp[0] = 0.0f;
    yp[0] = 0.0f;
    for (int b = 0; b < n_bins; ++b) {
        wp[b + 1] = wp[b] + wb[b];
        yp[b + 1] = yp[b] + yb[b];
    }
    const float wt = wp[n_bins];
    if (wt <= 0.0f) return 0.0f;
    float best = 0.0f;
    for (int64_t d = nb_min; d <= nb_max; ++d) {
        for (int64_t p0 = 0; p0 < n_bins; ++p0) {
            int64_t end = p0 + d;
            float r, s;
            if (end <= n_bins) {
                r = wp[end] - wp[p0];
                s = yp[end] - yp[p0];
            } else {
                int64_t wrap = end - n_bins;
                r = (wp[n_bins] - wp[p0]) + wp[wrap];
                s = (yp[n_bins] - yp[p0]) + yp[wrap];
            }
            float rf = r / wt;
            if (rf > 0.0f && rf < 1.0f) {
                float num = s * s;
                float den = r * (wt - r);
                float v = num / den;
                if (v > best) best = v;
            }
        }
    }
    return best;
}

int n_threads_or(int requested) {
#ifdef _OPENMP
    return requested > 0 ? requested : omp_get_max_threads();
#else
    (void)requested;
    return 1;
#endif
}

}  // namespace

extern "C" {

// Exposed for tests: the fold primitive (rust/src/fold.rs).
float or_fold_time(float t, float period, float period_dt) {
    return fold_time(t, period, pdt_correction(period, period_dt));
}

void or_fold_batch(const float* t, int64_t n, float period, float period_dt, float* out) {
    float k = pdt_correction(period, period_dt);
    for (int64_t i = 0; i < n; ++i) out[i] = fold_time(t[i], period, k);
}

// All batched entry points take CSR-packed curves: curve c owns points
// [off[c], off[c+1]) of t/y(/e).  Output is out[c][period][pdt] (lib.rs:85).

int or_ce_batched(int64_t n_curves, const int64_t* off, const float* t, const float* m,
                  const float* periods, int64_t np_, const float* pdts, int64_t nd,
                  int n_phase, int n_mag, int p_ext, int m_ext, float* out, int threads) {
    const int64_t T = np_ * nd;
    const int64_t total = n_curves * T;
#pragma omp parallel num_threads(n_threads_or(threads))
    {
        std::vector<uint32_t> hist(static_cast<size_t>(n_phase) * n_mag);
#pragma omp for schedule(dynamic, 16)
        for (int64_t g = 0; g < total; ++g) {
            int64_t c = g / T, tr = g % T;
            int64_t a = off[c], n = off[c + 1] - off[c];
            out[g] = ce_trial(t + a, m + a, n, periods[tr / nd], pdts[tr % nd], n_phase, n_mag,
                              p_ext, m_ext, hist.data());
        }
    }
    return 0;
}

int or_aov_batched(int64_t n_curves, const int64_t* off, const float* t, const float* m,
                   const float* periods, int64_t np_, const float* pdts, int64_t nd,
                   int n_bins, int ext, float* out, int threads) {
    const int64_t T = np_ * nd;
    const int64_t total = n_curves * T;
    std::vector<float> avg(n_curves);
    for (int64_t c = 0; c < n_curves; ++c) {
        float s = 0.0f;
        for (int64_t i = off[c]; i < off[c + 1]; ++i) s += m[i];
        avg[c] = s / static_cast<float>(off[c + 1] - off[c]);
    }
#pragma omp parallel num_threads(n_threads_or(threads))
    {
        std::vector<uint32_t> cnt(n_bins);
        std::vector<float> sum(n_bins), sq(n_bins);
#pragma omp for schedule(dynamic, 16)
        for (int64_t g = 0; g < total; ++g) {
            int64_t c = g / T, tr = g % T;
            int64_t a = off[c], n = off[c + 1] - off[c];
            out[g] = aov_trial(t + a, m + a, n, avg[c], periods[tr / nd], pdts[tr % nd], n_bins,
                               ext, cnt.data(), sum.data(), sq.data());
        }
    }
    return 0;
}

int or_ls_batched(int64_t n_curves, const int64_t* off, const float* t, const float* y,
                  const float* periods, int64_t np_, const float* pdts, int64_t nd, float* out,
                  int threads) {
    const int64_t T = np_ * nd;
    const int64_t total = n_curves * T;
#pragma omp parallel for schedule(dynamic, 16) num_threads(n_threads_or(threads))
    for (int64_t g = 0; g < total; ++g) {
        int64_t c = g / T, tr = g % T;
        int64_t a = off[c], n = off[c + 1] - off[c];
        out[g] = ls_trial(t + a, y + a, n, periods[tr / nd], pdts[tr % nd]);
    }
    return 0;
}

int or_fpw_batched(int64_t n_curves, const int64_t* off, const float* t, const float* y,
                   const float* e, const float* periods, int64_t np_, const float* pdts,
                   int64_t nd, int n_bins, float* out, int threads) {
    const int64_t T = np_ * nd;
    const int64_t total = n_curves * T;
    const int64_t npts = off[n_curves];
    std::vector<float> w(npts), wy(npts);
    for (int64_t i = 0; i < npts; ++i) {
        float ee = e[i] * e[i];
        w[i] = 1.0f / ee;
        wy[i] = w[i] * y[i];
    }
#pragma omp parallel num_threads(n_threads_or(threads))
    {
        std::vector<float> vw(n_bins), vy(n_bins);
#pragma omp for schedule(dynamic, 16)
        for (int64_t g = 0; g < total; ++g) {
            int64_t c = g / T, tr = g % T;
            int64_t a = off[c], n = off[c + 1] - off[c];
            out[g] = fpw_trial(t + a, w.data() + a, wy.data() + a, n, periods[tr / nd],
                               pdts[tr % nd], n_bins, vw.data(), vy.data());
        }
    }
    return 0;
}

// Per-curve BLS preprocessing (rust/src/bls.rs:34-51), shared with tests.
void or_bls_prepare(const float* y, const float* e, int64_t n, float* w, float* yw,
                    float* mean_out) {
    float total_w = 0.0f;
    for (int64_t i = 0; i < n; ++i) {
        float ee = e[i] * e[i];
        w[i] = 1.0f / ee;
        total_w += w[i];
    }
    float mean = 0.0f;
    if (total_w > 0.0f) {
        float s = 0.0f;
        for (int64_t i = 0; i < n; ++i) s += w[i] * y[i];
        mean = s / total_w;
    }
    for (int64_t i = 0; i < n; ++i) yw[i] = w[i] * (y[i] - mean);
    if (mean_out) *mean_out = mean;
}

void or_bls_duration_range(int n_bins, float qmin, float qmax, int64_t* lo, int64_t* hi) {
    const float nbf = static_cast<float>(n_bins);
    uint64_t a = f32_to_usize(std::floor(qmin * nbf));
    uint64_t b = f32_to_usize(std::ceil(qmax * nbf));
    // A saturated a (qmin * n_bins >= 2^64) leaves Rust's nb_min..=nb_max
    // empty; n_bins is above hi and keeps that range empty in int64.
    a = std::min<uint64_t>(a, static_cast<uint64_t>(n_bins));
    *lo = static_cast<int64_t>(std::max<uint64_t>(a, 1));
    *hi = static_cast<int64_t>(std::min<uint64_t>(b, static_cast<uint64_t>(n_bins - 1)));
}

int or_bls_batched(int64_t n_curves, const int64_t* off, const float* t, const float* y,
                   const float* e, const float* periods, int64_t np_, const float* pdts,
                   int64_t nd, int n_bins, float qmin, float qmax, float* out, int threads) {
    const int64_t T = np_ * nd;
    const int64_t total = n_curves * T;
    const int64_t npts = off[n_curves];
    std::vector<float> w(npts), yw(npts);
    for (int64_t c = 0; c < n_curves; ++c) {
        int64_t a = off[c];
        or_bls_prepare(y + a, e + a, off[c + 1] - a, w.data() + a, yw.data() + a, nullptr);
    }
    int64_t lo, hi;
    or_bls_duration_range(n_bins, qmin, qmax, &lo, &hi);
#pragma omp parallel num_threads(n_threads_or(threads))
    {
        std::vector<float> wb(n_bins), yb(n_bins), wp(n_bins + 1), yp(n_bins + 1);
#pragma omp for schedule(dynamic, 16)
        for (int64_t g = 0; g < total; ++g) {
            int64_t c = g / T, tr = g % T;
            int64_t a = off[c], n = off[c + 1] - off[c];
            out[g] = bls_trial(t + a, w.data() + a, yw.data() + a, n, periods[tr / nd],
                               pdts[tr % nd], n_bins, lo, hi, wb.data(), yb.data(), wp.data(),
                               yp.data());
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Peak finder: rust/src/peaks.rs:18-238 (streaming greedy top-N of strict
// local extrema with a minimum distance), single chunk + finalize.
// Returns the number of peaks written; idx/val have room for n_peaks.
// ---------------------------------------------------------------------------
int64_t or_find_top_peaks(const float* data, int64_t n, int64_t n_peaks, int64_t min_distance,
                          int use_max, int64_t* idx, float* val) {
    const float worst = use_max ? -std::numeric_limits<float>::infinity()
                                : std::numeric_limits<float>::infinity();
    auto better = [use_max](float a, float b) { return use_max ? (a > b) : (a < b); };
    std::vector<int64_t> li(n_peaks > 0 ? n_peaks : 0, 0);
    std::vector<float> lv(n_peaks > 0 ? n_peaks : 0, worst);
    int64_t count = 0;
    int64_t cand_i = -1;
    float cand_v = worst;
    bool has_cand = false;
    int64_t since = 0;

    auto threshold = [&]() { return count < n_peaks ? worst : lv[count - 1]; };
    auto insert = [&](int64_t i, float v) {
        int64_t pos;
        if (count < n_peaks) {
            pos = count++;
        } else if (count > 0 && better(v, lv[count - 1])) {
            pos = count - 1;
        } else {
            return;
        }
        li[pos] = i;
        lv[pos] = v;
        while (pos > 0 && better(lv[pos], lv[pos - 1])) {
            std::swap(lv[pos], lv[pos - 1]);
            std::swap(li[pos], li[pos - 1]);
            --pos;
        }
    };
    auto promote = [&]() {
        if (has_cand && better(cand_v, threshold())) insert(cand_i, cand_v);
        has_cand = false;
        since = 0;
    };
    // Interior points only: endpoints are never peaks (peaks.rs:186-188).
    for (int64_t i = 1; i + 1 < n; ++i) {
        float v = data[i], l = data[i - 1], r = data[i + 1];
        bool ext = use_max ? (v > l && v > r) : (v < l && v < r);
        if (ext && better(v, threshold())) {
            if (!has_cand || better(v, cand_v)) {
                cand_i = i;
                cand_v = v;
                has_cand = true;
                since = 0;
            }
        }
        if (has_cand) {
            ++since;
            if (since >= min_distance) promote();
        }
    }
    promote();
    for (int64_t k = 0; k < count; ++k) {
        idx[k] = li[k];
        val[k] = lv[k];
    }
    return count;
}

// Batched peaks over out[c][T] (lib.rs:509-555 semantics: idx=-1, val=0 padding).
void or_find_top_peaks_batched(const float* data, int64_t n_curves, int64_t T, int64_t n_peaks,
                               int64_t min_distance, int use_max, int64_t* idx, float* val) {
    for (int64_t c = 0; c < n_curves; ++c) {
        int64_t* ci = idx + c * n_peaks;
        float* cv = val + c * n_peaks;
        for (int64_t k = 0; k < n_peaks; ++k) { ci[k] = -1; cv[k] = 0.0f; }
        or_find_top_peaks(data + c * T, T, n_peaks, min_distance, use_max, ci, cv);
    }
}

int or_max_threads(void) { return n_threads_or(0); }

}  // extern "C"
