// CPU oracle: Viterbi Narrowband (rust/src/vn.rs).
//
// TEST INFRASTRUCTURE ONLY, like pf_oracle.cpp: a restatement of the Rust
// CPU backend for checking the HIP path; only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load it.
//
// Operation by operation as vn.rs, built with -ffp-contract=off; ln is the
// C library's logf, as Rust std calls it on Linux.
#include <cmath>
#include <cstdint>
#include <vector>

#include <omp.h>

namespace {

inline uint64_t f32_to_usize(float x) {
    if (!(x > 0.0f)) return 0;
    if (x >= 18446744073709551616.0f) return UINT64_MAX;
    return static_cast<uint64_t>(x);
}

inline float pdt_correction(float period, float period_dt) {
    return (period_dt / period) / 2.0f;
}

inline float fold_time(float t, float period, float pdt_corr) {
    const double td = static_cast<double>(t);
    const double k = static_cast<double>(pdt_corr);
    const double r = (td - (k * td) * td) / static_cast<double>(period);
    return static_cast<float>(std::fabs(r - std::floor(r)));
}

// vn.rs:17-150 (circular_viterbi_concentration); hist[phi * n_mag + m].
float viterbi_concentration(const float* hist, int np_, int nm, int bw, int margin,
                            std::vector<float>& emit, std::vector<int>& back,
                            std::vector<int>& best_path) {
    const float eps = 1e-10f;
    const float neg_inf = -INFINITY;
    for (int i = 0; i < np_ * nm; ++i) {
        const float x = hist[i] + eps;
        emit[i] = logf(x);
    }
    float best_total = neg_inf;
    std::vector<float> v_prev(nm), v_curr(nm);
    for (int start = 0; start < nm; ++start) {
        for (int m = 0; m < nm; ++m) v_prev[m] = v_curr[m] = neg_inf;
        v_prev[start] = emit[start];
        for (int phi = 1; phi < np_; ++phi) {
            for (int m = 0; m < nm; ++m) {
                float best_prev = neg_inf;
                int best_prev_m = 0;
                const int lo = m >= bw ? m - bw : 0;
                const int hi = m + bw < nm ? m + bw : nm - 1;
                for (int mp = lo; mp <= hi; ++mp)
                    if (v_prev[mp] > best_prev) {
                        best_prev = v_prev[mp];
                        best_prev_m = mp;
                    }
                v_curr[m] = emit[phi * nm + m] + best_prev;
                back[phi * nm + m] = best_prev_m;
            }
            std::swap(v_prev, v_curr);
            for (int m = 0; m < nm; ++m) v_curr[m] = neg_inf;
        }
        const float total = v_prev[start];
        if (total > best_total) {
            best_total = total;
            best_path[np_ - 1] = start;
            for (int phi = np_ - 1; phi >= 1; --phi)
                best_path[phi - 1] = back[phi * nm + best_path[phi]];
        }
    }
    if (best_total == neg_inf) return 0.0f;
    float sum_frac = 0.0f;
    uint32_t n_cols = 0;
    for (int phi = 0; phi < np_; ++phi) {
        float col_mass = 0.0f;
        for (int m = 0; m < nm; ++m) col_mass += hist[phi * nm + m];
        if (col_mass <= 0.0f) continue;
        const int pm = best_path[phi];
        const int lo = pm >= margin ? pm - margin : 0;
        const int hi = pm + margin < nm ? pm + margin : nm - 1;
        float path_mass = 0.0f;
        for (int m = lo; m <= hi; ++m) path_mass += hist[phi * nm + m];
        sum_frac += path_mass / col_mass;
        n_cols += 1;
    }
    if (n_cols == 0) return 0.0f;
    const float avg = sum_frac / static_cast<float>(n_cols);
    const int band_i = (2 * margin + 1) < nm ? (2 * margin + 1) : nm;
    const float band = static_cast<float>(band_i);
    const float baseline = band / static_cast<float>(nm);
    float r = (avg - baseline) / (1.0f - baseline);
    r = std::isnan(r) ? 0.0f : (r > 0.0f ? r : 0.0f);  // f32::max(0.0): NaN -> 0.0
    r = r < 1.0f ? r : 1.0f;                             // f32::min(1.0)
    return r;
}

// vn.rs:160-219 for one trial.
float vn_trial(const float* t, const float* m, int64_t n, float period, float pdt, int np_,
               int nm, int pext, int mext, int bw, int margin, std::vector<uint32_t>& hist,
               std::vector<float>& hf, std::vector<float>& emit, std::vector<int>& back,
               std::vector<int>& path) {
    const float pbs = 1.0f / static_cast<float>(np_);
    const float mbs = 1.0f / static_cast<float>(nm);
    const float norm = static_cast<float>(static_cast<uint64_t>(n) * pext * mext);
    const float k = pdt_correction(period, pdt);
    for (auto& h : hist) h = 0;
    for (int64_t i = 0; i < n; ++i) {
        const float folded = fold_time(t[i], period, k);
        const uint64_t pb0 = f32_to_usize(folded / pbs);
        const uint64_t mb0 = f32_to_usize(m[i] / mbs);
        for (int po = 0; po < pext; ++po)
            for (int mo = 0; mo < mext; ++mo) {
                const uint64_t pb = (pb0 + po) % static_cast<uint64_t>(np_);
                const uint64_t mb = (mb0 + mo) % static_cast<uint64_t>(nm);
                hist[pb * nm + mb] += 1;
            }
    }
    for (int i = 0; i < np_ * nm; ++i) hf[i] = static_cast<float>(hist[i]) / norm;
    return viterbi_concentration(hf.data(), np_, nm, bw, margin, emit, back, path);
}

int n_threads(int requested) { return requested > 0 ? requested : omp_get_max_threads(); }

}  // namespace

extern "C" {

// calc_vn (vn.rs:156-219) over CSR curves: out[c][period][pdt].
int or_vn_batched(int64_t n_curves, const int64_t* off, const float* t, const float* m,
                  const float* periods, int64_t np_, const float* pdts, int64_t nd, int n_phase,
                  int n_mag, int phase_ext, int mag_ext, int bandwidth, int margin, float* out,
                  int threads) {
    if (n_phase < 1 || n_mag < 1 || phase_ext < 1 || mag_ext < 1 || bandwidth < 0 || margin < 0)
        return -1;
    const int64_t T = np_ * nd;
    const int64_t total = n_curves * T;
#pragma omp parallel num_threads(n_threads(threads))
    {
        const int cells = n_phase * n_mag;
        std::vector<uint32_t> hist(cells);
        std::vector<float> hf(cells), emit(cells);
        std::vector<int> back(cells), path(n_phase);
#pragma omp for schedule(dynamic, 16)
        for (int64_t g = 0; g < total; ++g) {
            const int64_t c = g / T, tr = g % T;
            const int64_t a = off[c], n = off[c + 1] - off[c];
            out[g] = vn_trial(t + a, m + a, n, periods[tr / nd], pdts[tr % nd], n_phase, n_mag,
                              phase_ext, mag_ext, bandwidth, margin, hist, hf, emit, back, path);
        }
    }
    return 0;
}

// The concentration score of one histogram (vn.rs:17-150), for known answers.
float or_vn_concentration(const float* hist, int n_phase, int n_mag, int bandwidth, int margin) {
    const int cells = n_phase * n_mag;
    std::vector<float> emit(cells);
    std::vector<int> back(cells), path(n_phase);
    return viterbi_concentration(hist, n_phase, n_mag, bandwidth, margin, emit, back, path);
}

}  // extern "C"
