// CPU oracle: Multi-Harmonic Fourier (rust/src/mhf.rs, with cholesky_solve
// of rust/src/fourier.rs:23-61) and the Fourier decomposition features
// (rust/src/fourier.rs:68-230).
//
// TEST INFRASTRUCTURE ONLY, like pf_oracle.cpp: a restatement of the Rust
// CPU backend for checking the HIP path; only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load it.
//
// Operation by operation as mhf.rs, in f64, built with -ffp-contract=off;
// sin/cos are the C library's, as Rust std calls them on Linux.
#include <cmath>
#include <cstdint>
#include <vector>

#include <omp.h>

namespace {

constexpr int kMaxK = 5;
constexpr int kMaxParams = 2 + 2 * kMaxK;

inline float pdt_correction(float period, float period_dt) {
    return (period_dt / period) / 2.0f;
}

inline float fold_time(float t, float period, float pdt_corr) {
    const double td = static_cast<double>(t);
    const double k = static_cast<double>(pdt_corr);
    const double r = (td - (k * td) * td) / static_cast<double>(period);
    return static_cast<float>(std::fabs(r - std::floor(r)));
}

// fourier.rs:23-61; a is n x n row-major (lower triangle read), b -> x.
bool cholesky_solve(double* a, double* b, int n) {
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) {
            double s = a[i * n + j];
            for (int k = 0; k < j; ++k) s -= a[i * n + k] * a[j * n + k];
            if (i == j) {
                if (s <= 0.0) return false;
                a[i * n + j] = std::sqrt(s);
            } else {
                a[i * n + j] = s / a[j * n + j];
            }
        }
    for (int i = 0; i < n; ++i) {
        double s = b[i];
        for (int k = 0; k < i; ++k) s -= a[i * n + k] * b[k];
        b[i] = s / a[i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = b[i];
        for (int k = i + 1; k < n; ++k) s -= a[k * n + i] * b[k];
        b[i] = s / a[i * n + i];
    }
    return true;
}

struct Result {
    double delta[kMaxK + 1];
    double best;
    int best_k;
};

// mhf.rs:75-226 (eval_mhf_single_period).
Result eval(const float* times, const double* y, const double* w, int64_t n, double ln_n,
            float period, float period_dt, int max_k, std::vector<double>& xc) {
    Result r{};
    const float pdt_corr = pdt_correction(period, period_dt);
    if (period <= 0.0f || !std::isfinite(period)) return r;
    float tmin = INFINITY;
    for (int64_t i = 0; i < n; ++i) tmin = std::fmin(tmin, times[i]);  // f32::min: NaN ignored
    const int np_full = 2 + 2 * max_k;
    auto X = [&](int c, int64_t d) -> double& { return xc[static_cast<size_t>(c) * n + d]; };
    for (int64_t i = 0; i < n; ++i) {
        const float phase = fold_time(times[i], period, pdt_corr);
        const double phi = 2.0 * M_PI * static_cast<double>(phase);
        X(0, i) = 1.0;
        X(1, i) = static_cast<double>(times[i] - tmin);
        if (max_k >= 1) {
            const double s1 = std::sin(phi), c1 = std::cos(phi);
            X(2, i) = c1;
            X(3, i) = s1;
            double cp2 = 1.0, sp2 = 0.0, cp1 = c1, sp1 = s1;
            const double two_c1 = 2.0 * c1;
            for (int h = 2; h <= max_k; ++h) {
                const double ck = two_c1 * cp1 - cp2;
                const double sk = two_c1 * sp1 - sp2;
                X(2 * h, i) = ck;
                X(2 * h + 1, i) = sk;
                cp2 = cp1;
                sp2 = sp1;
                cp1 = ck;
                sp1 = sk;
            }
        }
    }
    double ata[kMaxParams * kMaxParams], atb[kMaxParams];
    for (int cj = 0; cj < np_full; ++cj) {
        for (int ci = 0; ci <= cj; ++ci) {
            double s = 0.0;
            for (int64_t d = 0; d < n; ++d) s += w[d] * X(ci, d) * X(cj, d);
            ata[cj * np_full + ci] = s;
            ata[ci * np_full + cj] = s;
        }
        double s = 0.0;
        for (int64_t d = 0; d < n; ++d) s += w[d] * X(cj, d) * y[d];
        atb[cj] = s;
    }
    double chi2[kMaxK + 1];
    bool ok[kMaxK + 1];
    for (int k = 0; k <= kMaxK; ++k) {
        chi2[k] = INFINITY;
        ok[k] = false;
    }
    for (int k = 0; k <= max_k; ++k) {
        const int npk = 2 + 2 * k;
        if (n <= npk) continue;
        double sub[kMaxParams * kMaxParams], sol[kMaxParams];
        for (int row = 0; row < npk; ++row) {
            for (int col = 0; col < npk; ++col) sub[row * npk + col] = ata[row * np_full + col];
            sol[row] = atb[row];
        }
        if (!cholesky_solve(sub, sol, npk)) continue;
        ok[k] = true;
        double c2 = 0.0;
        for (int64_t d = 0; d < n; ++d) {
            double pred = 0.0;
            for (int j = 0; j < npk; ++j) pred += sol[j] * X(j, d);
            const double resid = y[d] - pred;
            c2 += w[d] * resid * resid;
        }
        chi2[k] = c2;
    }
    if (!ok[0]) return r;
    const double bic_flat = chi2[0] + ln_n * 2.0;
    double best = 0.0;
    int best_k = 0;
    for (int k = 0; k <= max_k; ++k) {
        if (!ok[k]) {
            r.delta[k] = 0.0;
            continue;
        }
        const double npk = static_cast<double>(2 + 2 * k);
        const double bic_k = chi2[k] + ln_n * npk;
        double delta = bic_flat - bic_k;
        delta = std::isnan(delta) ? 0.0 : std::fmax(delta, 0.0);  // f64::max(0.0)
        r.delta[k] = delta;
        if (delta > best) {
            best = delta;
            best_k = k;
        }
    }
    r.best = best;
    r.best_k = best_k;
    return r;
}

void prepare(const float* m, const float* e, int64_t n, std::vector<double>& y,
             std::vector<double>& w) {
    y.resize(n);
    w.resize(n);
    for (int64_t i = 0; i < n; ++i) {
        y[i] = static_cast<double>(m[i]);
        const double e64 = static_cast<double>(e ? e[i] : 1.0f);
        w[i] = e64 > 0.0 ? 1.0 / (e64 * e64) : 0.0;
    }
}

int n_threads(int requested) { return requested > 0 ? requested : omp_get_max_threads(); }

// fourier.rs:68-230 (calc_fourier_single): 14 features at one period.
void fourier_single(const float* times, const float* mags, const float* errs, int64_t n,
                    float period, float* out, std::vector<double>& xc) {
    constexpr int kF = 14;
    for (int j = 0; j < kF; ++j) out[j] = NAN;
    if (n < 3 || period <= 0.0f || !std::isfinite(period)) return;
    float tmin = INFINITY;
    for (int64_t i = 0; i < n; ++i) tmin = std::fmin(tmin, times[i]);
    std::vector<double> w(n), y(n);
    xc.assign(static_cast<size_t>(kMaxParams) * n, 0.0);
    auto X = [&](int c, int64_t d) -> double& { return xc[static_cast<size_t>(c) * n + d]; };
    for (int64_t i = 0; i < n; ++i) {
        const double e64 = static_cast<double>(errs[i]);
        w[i] = e64 > 0.0 ? 1.0 / (e64 * e64) : 0.0;
        y[i] = static_cast<double>(mags[i]);
        X(0, i) = 1.0;
        X(1, i) = static_cast<double>(times[i] - tmin);
        const double phi = 2.0 * M_PI * static_cast<double>(times[i]) / static_cast<double>(period);
        for (int h = 1; h <= kMaxK; ++h) {
            X(2 * h, i) = std::cos(static_cast<double>(h) * phi);
            X(2 * h + 1, i) = std::sin(static_cast<double>(h) * phi);
        }
    }
    const double ln_n = std::log(static_cast<double>(n));
    double chi2[kMaxK + 1];
    double params[kMaxK + 1][kMaxParams] = {};
    bool ok[kMaxK + 1] = {};
    for (int k = 0; k <= kMaxK; ++k) {
        const int npk = 2 + 2 * k;
        double a[kMaxParams * kMaxParams], b[kMaxParams];
        for (int cj = 0; cj < npk; ++cj) {
            for (int ci = 0; ci <= cj; ++ci) {
                double s = 0.0;
                for (int64_t d = 0; d < n; ++d) s += w[d] * X(ci, d) * X(cj, d);
                a[cj * npk + ci] = s;
                a[ci * npk + cj] = s;
            }
            double s = 0.0;
            for (int64_t d = 0; d < n; ++d) s += w[d] * X(cj, d) * y[d];
            b[cj] = s;
        }
        if (!cholesky_solve(a, b, npk)) {
            chi2[k] = INFINITY;
            continue;
        }
        ok[k] = true;
        for (int j = 0; j < npk; ++j) params[k][j] = b[j];
        double c2 = 0.0;
        for (int64_t d = 0; d < n; ++d) {
            double pred = 0.0;
            for (int j = 0; j < npk; ++j) pred += b[j] * X(j, d);
            const double resid = y[d] - pred;
            c2 += w[d] * resid * resid;
        }
        chi2[k] = c2;
    }
    if (!ok[0]) return;
    int best = 0;
    double best_bic = INFINITY;
    for (int k = 0; k <= kMaxK; ++k) {
        if (!ok[k]) continue;
        const double bic = chi2[k] + ln_n * static_cast<double>(2 + 2 * k);
        if (bic < best_bic) {
            best_bic = bic;
            best = k;
        }
    }
    const double power = chi2[0] > 0.0 ? (chi2[0] - chi2[best]) / chi2[0] : 0.0;
    out[0] = static_cast<float>(power);
    out[1] = static_cast<float>(best_bic);
    for (int j = 0; j < kMaxParams; ++j) out[2 + j] = static_cast<float>(params[best][j]);
}

}  // namespace

extern "C" {

// calc_fourier_batch (fourier.rs:236-253) over CSR curves: out[c][14].
int or_fourier_batched(int64_t n_curves, const int64_t* off, const float* t, const float* m,
                       const float* e, const float* periods, float* out, int threads) {
#pragma omp parallel num_threads(n_threads(threads))
    {
        std::vector<double> xc;
#pragma omp for schedule(dynamic, 1)
        for (int64_t c = 0; c < n_curves; ++c) {
            const int64_t a = off[c];
            fourier_single(t + a, m + a, e + a, off[c + 1] - a, periods[c], out + c * 14, xc);
        }
    }
    return 0;
}

// cholesky_solve (fourier.rs:23-61) on one system, for its known answers:
// a[n*n] row-major, b[n] -> x; returns 1 on success.
int or_cholesky_solve(double* a, double* b, int n) { return cholesky_solve(a, b, n) ? 1 : 0; }

// calc_mhf (mhf.rs:234-291) over CSR curves: out[c][period][pdt] = best dBIC.
int or_mhf_batched(int64_t n_curves, const int64_t* off, const float* t, const float* m,
                   const float* e, const float* periods, int64_t np_, const float* pdts,
                   int64_t nd, int max_harmonics, float* out, int threads) {
    if (max_harmonics < 0) return -1;
    const int max_k = max_harmonics < kMaxK ? max_harmonics : kMaxK;
    const int64_t T = np_ * nd;
#pragma omp parallel num_threads(n_threads(threads))
    {
        std::vector<double> y, w, xc;
#pragma omp for schedule(dynamic, 1)
        for (int64_t c = 0; c < n_curves; ++c) {
            const int64_t a = off[c], n = off[c + 1] - off[c];
            float* o = out + c * T;
            if (n < 3) {
                for (int64_t g = 0; g < T; ++g) o[g] = 0.0f;
                continue;
            }
            prepare(m + a, e ? e + a : nullptr, n, y, w);
            xc.assign(static_cast<size_t>(kMaxParams) * n, 0.0);
            const double ln_n = std::log(static_cast<double>(n));
            for (int64_t g = 0; g < T; ++g)
                o[g] = static_cast<float>(eval(t + a, y.data(), w.data(), n, ln_n,
                                               periods[g / nd], pdts[g % nd], max_k, xc)
                                              .best);
        }
    }
    return 0;
}

// calc_mhf_per_k (mhf.rs:299-334) at one (period, period_dt) per curve:
// out[c][max_k + 2] = dBIC for k = 0..max_k, then best_k.
int or_mhf_per_k(int64_t n_curves, const int64_t* off, const float* t, const float* m,
                 const float* e, const float* periods, const float* pdts, int max_harmonics,
                 float* out, int threads) {
    if (max_harmonics < 0) return -1;
    const int max_k = max_harmonics < kMaxK ? max_harmonics : kMaxK;
    const int width = max_k + 2;
#pragma omp parallel num_threads(n_threads(threads))
    {
        std::vector<double> y, w, xc;
#pragma omp for schedule(dynamic, 1)
        for (int64_t c = 0; c < n_curves; ++c) {
            const int64_t a = off[c], n = off[c + 1] - off[c];
            float* o = out + c * width;
            for (int j = 0; j < width; ++j) o[j] = 0.0f;
            if (n < 3) continue;
            prepare(m + a, e ? e + a : nullptr, n, y, w);
            xc.assign(static_cast<size_t>(kMaxParams) * n, 0.0);
            const double ln_n = std::log(static_cast<double>(n));
            const Result r =
                eval(t + a, y.data(), w.data(), n, ln_n, periods[c], pdts[c], max_k, xc);
            for (int k = 0; k <= max_k; ++k) o[k] = static_cast<float>(r.delta[k]);
            o[max_k + 1] = static_cast<float>(r.best_k);
        }
    }
    return 0;
}

}  // extern "C"
