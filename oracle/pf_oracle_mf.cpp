// CPU oracle: Matched Filter (rust/src/mf.rs).
//
// TEST INFRASTRUCTURE ONLY, like pf_oracle.cpp: a restatement of the Rust
// CPU backend for checking the HIP path; only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load it.
//
// Operation by operation as mf.rs, built with -ffp-contract=off.  Rust's
// `iter().sum::<f32>()` folds from -0.0 (the float Sum neutral element of
// current Rust), serially; f32 sin is the platform libm's sinf, as Rust std
// calls it on Linux.
#include <cmath>
#include <cstdint>
#include <vector>

#include <omp.h>

namespace {

inline uint64_t f32_to_usize(float x) {
    if (!(x > 0.0f)) return 0;  // NaN and negatives saturate to 0
    if (x >= 18446744073709551616.0f) return UINT64_MAX;
    return static_cast<uint64_t>(x);
}

inline float pdt_correction(float period, float period_dt) {
    return (period_dt / period) / 2.0f;  // rust/src/fold.rs:13-15
}

inline float fold_time(float t, float period, float pdt_corr) {  // fold.rs:22-26
    const double td = static_cast<double>(t);
    const double k = static_cast<double>(pdt_corr);
    const double r = (td - (k * td) * td) / static_cast<double>(period);
    return static_cast<float>(std::fabs(r - std::floor(r)));
}

float rust_sum(const float* x, int n) {
    float s = -0.0f;
    for (int i = 0; i < n; ++i) s = s + x[i];
    return s;
}

// mf.rs:29-38
void normalize_template(std::vector<float>& t) {
    const int n_i = static_cast<int>(t.size());
    const float n = static_cast<float>(n_i);
    const float mean = rust_sum(t.data(), n_i) / n;
    for (float& x : t) x -= mean;
    float var = -0.0f;
    for (float x : t) var = var + x * x;
    var = var / n;
    const float sd = std::sqrt(var);
    if (sd > 1e-12f)
        for (float& x : t) x /= sd;
}

enum Kind { kSaw = 0, kSin = 1, kEcl = 2 };

// mf.rs:49-112: 5 sawtooth, 1 sinusoid, 8 eclipsing shapes.
void generate_templates(int nb, std::vector<std::vector<float>>& shapes, std::vector<int>& kinds) {
    const float nbf = static_cast<float>(nb);
    const float rise[5] = {0.10f, 0.20f, 0.35f, 0.50f, 0.65f};
    for (float rf : rise) {
        std::vector<float> t(nb, 0.0f);
        for (int k = 0; k < nb; ++k) {
            const float phase = (static_cast<float>(k) + 0.5f) / nbf;
            t[k] = phase < rf ? phase / rf : 1.0f - (phase - rf) / (1.0f - rf);
        }
        normalize_template(t);
        shapes.push_back(t);
        kinds.push_back(kSaw);
    }
    {
        const float two_pi = 2.0f * static_cast<float>(M_PI);  // 2.0 * f32::consts::PI
        std::vector<float> t(nb, 0.0f);
        for (int k = 0; k < nb; ++k) {
            const float phase = (static_cast<float>(k) + 0.5f) / nbf;
            t[k] = sinf(two_pi * phase);
        }
        normalize_template(t);
        shapes.push_back(t);
        kinds.push_back(kSin);
    }
    const float widths[4] = {0.05f, 0.10f, 0.15f, 0.20f};
    const bool has_sec[2] = {false, true};
    const float sec_frac[2] = {0.0f, 0.5f};
    for (float dw : widths) {
        for (int o = 0; o < 2; ++o) {
            std::vector<float> t(nb, 1.0f);
            const float half = dw / 2.0f;
            for (int k = 0; k < nb; ++k) {
                const float phase = (static_cast<float>(k) + 0.5f) / nbf;
                const float dist_p = phase < 0.5f ? phase : 1.0f - phase;
                if (dist_p < half) t[k] = 0.0f;
                if (has_sec[o]) {
                    const float dist_s = std::fabs(phase - 0.5f);
                    if (dist_s < half) t[k] = 1.0f - sec_frac[o];
                }
            }
            normalize_template(t);
            shapes.push_back(t);
            kinds.push_back(kEcl);
        }
    }
}

// mf.rs:124-153
float circular_pearson_max(const float* profile, const float* tmpl, int nb) {
    const float n = static_cast<float>(nb);
    const float p_mean = rust_sum(profile, nb) / n;
    float p_var = -0.0f;
    for (int k = 0; k < nb; ++k) p_var = p_var + (profile[k] - p_mean) * (profile[k] - p_mean);
    if (p_var < 1e-20f) return 0.0f;
    const float p_std = std::sqrt(p_var);
    const float t_std = std::sqrt(n);
    const float denom = p_std * t_std;
    if (denom < 1e-20f) return 0.0f;
    float best = -1.0f;
    for (int shift = 0; shift < nb; ++shift) {
        float dot = 0.0f;
        for (int k = 0; k < nb; ++k) {
            const int tk = (k + shift) % nb;
            dot += (profile[k] - p_mean) * tmpl[tk];
        }
        const float r = dot / denom;
        if (r > best) best = r;
    }
    return best;
}

struct MfResult {
    float f[7];  // sawtooth, sinusoidal, eclipsing, R2, amp_snr, n_filled, combined
};

// mf.rs:172-330 (eval_mf_at_period).
MfResult eval_mf(const float* times, const float* mags, const float* ivar, const float* ivar_y,
                 int64_t n_data, float period, float pdt_corr, int nb,
                 const std::vector<std::vector<float>>& shapes, const std::vector<int>& kinds,
                 float* vt, float* yt, float* v2) {
    for (int k = 0; k < nb; ++k) vt[k] = yt[k] = v2[k] = 0.0f;
    const float nbf = static_cast<float>(nb);
    for (int64_t i = 0; i < n_data; ++i) {
        const float phase = fold_time(times[i], period, pdt_corr);
        const uint64_t bin =
            std::min<uint64_t>(f32_to_usize(phase * nbf), static_cast<uint64_t>(nb - 1));
        vt[bin] += ivar[i];
        yt[bin] += ivar_y[i];
        v2[bin] += ivar[i] * mags[i] * mags[i];
    }
    std::vector<float> profile(nb, 0.0f);
    uint32_t n_filled = 0;
    float total_wt = 0.0f, total_wy = 0.0f;
    for (int k = 0; k < nb; ++k) {
        if (vt[k] > 0.0f) {
            profile[k] = yt[k] / vt[k];
            n_filled += 1;
            total_wt += vt[k];
            total_wy += yt[k];
        }
    }
    const float coverage = static_cast<float>(n_filled) / nbf;
    MfResult r{};
    r.f[5] = static_cast<float>(n_filled);
    if (n_filled < 3) return r;
    const float grand_mean = total_wy / total_wt;
    float ss_total = 0.0f;
    for (int64_t i = 0; i < n_data; ++i) {
        const float d = mags[i] - grand_mean;
        ss_total += ivar[i] * d * d;
    }
    float ss_within = 0.0f;
    for (int k = 0; k < nb; ++k)
        if (vt[k] > 0.0f) ss_within += v2[k] - yt[k] * yt[k] / vt[k];
    float r2 = 0.0f;
    if (ss_total > 0.0f) {
        const float q = 1.0f - ss_within / ss_total;
        r2 = std::isnan(q) ? 0.0f : std::fmax(q, 0.0f);  // Rust f32::max: NaN -> other
    }
    float min_p = INFINITY, max_p = -INFINITY;
    for (int k = 0; k < nb; ++k)
        if (vt[k] > 0.0f) {
            if (profile[k] < min_p) min_p = profile[k];
            if (profile[k] > max_p) max_p = profile[k];
        }
    const float amplitude = max_p - min_p;
    float sum_rms = 0.0f;
    uint32_t n_rms = 0;
    for (int k = 0; k < nb; ++k)
        if (vt[k] > 0.0f) {
            const float wv = (v2[k] - yt[k] * yt[k] / vt[k]) / vt[k];
            if (wv > 0.0f) {
                sum_rms += std::sqrt(wv);
                n_rms += 1;
            }
        }
    const float mean_rms = n_rms > 0 ? sum_rms / static_cast<float>(n_rms) : 1.0f;
    const float amp_snr = mean_rms > 1e-12f ? amplitude / mean_rms : 0.0f;
    float best[3] = {0.0f, 0.0f, 0.0f};
    for (size_t s = 0; s < shapes.size(); ++s) {
        const float corr = circular_pearson_max(profile.data(), shapes[s].data(), nb);
        if (corr > best[kinds[s]]) best[kinds[s]] = corr;
    }
    // best_sawtooth.max(best_sinusoidal).max(best_eclipsing): Rust f32::max
    // returns the other operand when one is NaN (never the case here: the
    // bests start at 0 and only take values that compared greater).
    const float max_corr = std::fmax(std::fmax(best[0], best[1]), best[2]);
    r.f[0] = best[0];
    r.f[1] = best[1];
    r.f[2] = best[2];
    r.f[3] = r2;
    r.f[4] = amp_snr;
    r.f[6] = max_corr * r2 * coverage;
    return r;
}

int n_threads(int requested) {
    return requested > 0 ? requested : omp_get_max_threads();
}

}  // namespace

extern "C" {

// Templates as generated by mf.rs:49-112, shapes[14][nb] flattened; kinds[14]
// (0 sawtooth, 1 sinusoidal, 2 eclipsing).  Returns the number of shapes.
int or_mf_templates(int n_bins, float* shapes, int* kinds) {
    std::vector<std::vector<float>> sh;
    std::vector<int> kd;
    generate_templates(n_bins, sh, kd);
    for (size_t s = 0; s < sh.size(); ++s) {
        for (int k = 0; k < n_bins; ++k) shapes[s * n_bins + k] = sh[s][k];
        kinds[s] = kd[s];
    }
    return static_cast<int>(sh.size());
}

// mf.rs:124-153, exposed for the known-answer tests.
float or_mf_circular_pearson_max(const float* profile, const float* tmpl, int n_bins) {
    return circular_pearson_max(profile, tmpl, n_bins);
}

// calc_mf (mf.rs:344-393) over CSR curves: out[c][period][pdt] = combined.
int or_mf_batched(int64_t n_curves, const int64_t* off, const float* t, const float* y,
                  const float* e, const float* periods, int64_t np_, const float* pdts,
                  int64_t nd, int n_bins, float* out, int threads) {
    if (n_bins < 1) return -1;
    std::vector<std::vector<float>> shapes;
    std::vector<int> kinds;
    generate_templates(n_bins, shapes, kinds);
    const int64_t npts = off[n_curves];
    std::vector<float> ivar(npts), ivar_y(npts);
    for (int64_t i = 0; i < npts; ++i) {
        const float ei = e ? e[i] : 1.0f;
        ivar[i] = 1.0f / (ei * ei);
        ivar_y[i] = ivar[i] * y[i];
    }
    const int64_t T = np_ * nd;
    const int64_t total = n_curves * T;
#pragma omp parallel num_threads(n_threads(threads))
    {
        std::vector<float> vt(n_bins), yt(n_bins), v2(n_bins);
#pragma omp for schedule(dynamic, 16)
        for (int64_t g = 0; g < total; ++g) {
            const int64_t c = g / T, tr = g % T;
            const int64_t a = off[c], n = off[c + 1] - off[c];
            const float period = periods[tr / nd];
            const float k = pdt_correction(period, pdts[tr % nd]);
            out[g] = eval_mf(t + a, y + a, ivar.data() + a, ivar_y.data() + a, n, period, k,
                             n_bins, shapes, kinds, vt.data(), yt.data(), v2.data())
                         .f[6];
        }
    }
    return 0;
}

// calc_mf_features_batch (mf.rs:404-451): one period per curve, pdt 0;
// out[c][7].
int or_mf_features(int64_t n_curves, const int64_t* off, const float* t, const float* y,
                   const float* e, const float* periods, int n_bins, float* out, int threads) {
    if (n_bins < 1) return -1;
    std::vector<std::vector<float>> shapes;
    std::vector<int> kinds;
    generate_templates(n_bins, shapes, kinds);
#pragma omp parallel num_threads(n_threads(threads))
    {
        std::vector<float> vt(n_bins), yt(n_bins), v2(n_bins);
#pragma omp for schedule(dynamic, 1)
        for (int64_t c = 0; c < n_curves; ++c) {
            const int64_t a = off[c], n = off[c + 1] - off[c];
            std::vector<float> iv(n), ivy(n);
            for (int64_t i = 0; i < n; ++i) {
                const float ei = e[a + i];
                iv[i] = 1.0f / (ei * ei);
                ivy[i] = iv[i] * y[a + i];
            }
            const MfResult r = eval_mf(t + a, y + a, iv.data(), ivy.data(), n, periods[c], 0.0f,
                                       n_bins, shapes, kinds, vt.data(), yt.data(), v2.data());
            for (int f = 0; f < 7; ++f) out[c * 7 + f] = r.f[f];
        }
    }
    return 0;
}

}  // extern "C"
