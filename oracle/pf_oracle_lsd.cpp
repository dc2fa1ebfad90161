// CPU oracle for the differentiable Lomb-Scargle path (long double).
//
// TEST INFRASTRUCTURE ONLY (same rules as pf_oracle.cpp: loaded by tests/,
// __graft_entry__.smoke() and bench.py's CPU leg as the checker, never by the
// product).  The reference has no differentiable LS; what it fixes is the
// power, rust/src/ls.rs:15-79 (unweighted, no floating mean, SS = N - CC,
// tau-rotated, 0 when a denominator is 0).  This file restates:
//
//   or_lsd_tau_power  the reference's tau formula (ls.rs:41-74) at an
//                     arbitrary angular frequency w, with theta_i = w * t_i
//                     in long double instead of the f32 phase fold;
//   or_lsd_jet        P, dP/dw, d2P/dw2 and dP/dy_i from the quadratic form
//                     P = 1/2 b^T M^-1 b (b = sum y u, M = sum u u^T,
//                     u = (cos, sin)), on the raw times (no shift);
//   or_lsd_refine     safeguarded Newton on dP/dw inside a bracket, the
//                     algorithm the device refinement implements.
//
// Each is pinned by checks that do not reuse its own formula (tests/): the
// tau form against the f32 restatement of ls.rs at grid frequencies, the
// quadratic form against the tau form, every derivative against central
// finite differences of the tau form, the refinement against a golden-section
// search of the tau form.
#include <cmath>
#include <cstdint>

namespace {

using ld = long double;

struct Sums {
    ld C = 0, S = 0, A = 0, B = 0;     // sum y c, y s, c2, s2
    ld C1 = 0, S1 = 0, A1 = 0, B1 = 0;  // t-weighted
    ld C2 = 0, S2 = 0, A2 = 0, B2 = 0;  // t^2-weighted
};

Sums sums(const double* t, const double* y, int64_t n, ld w) {
    Sums q;
    for (int64_t i = 0; i < n; ++i) {
        const ld ti = t[i], yi = y[i];
        const ld th = w * ti;
        const ld c = cosl(th), s = sinl(th);
        const ld c2 = cosl(2 * th), s2 = sinl(2 * th);
        q.C += yi * c;
        q.S += yi * s;
        q.A += c2;
        q.B += s2;
        q.C1 += yi * ti * c;
        q.S1 += yi * ti * s;
        q.A1 += ti * c2;
        q.B1 += ti * s2;
        q.C2 += yi * ti * ti * c;
        q.S2 += yi * ti * ti * s;
        q.A2 += ti * ti * c2;
        q.B2 += ti * ti * s2;
    }
    return q;
}

struct JetLd {
    ld p, dp, d2p, x0, x1;
};

JetLd jet(const Sums& q, ld n) {
    JetLd j{0, 0, 0, 0, 0};
    const ld cc = (n + q.A) / 2, ss = (n - q.A) / 2, cs = q.B / 2;
    const ld det = cc * ss - cs * cs;
    if (!(det > 0)) return j;
    // x solves M x = b.
    const ld x0 = (ss * q.C - cs * q.S) / det;
    const ld x1 = (cc * q.S - cs * q.C) / det;
    j.x0 = x0;
    j.x1 = x1;
    j.p = (q.C * x0 + q.S * x1) / 2;
    // dM/dw: d(cc) = -B1, d(cs) = A1, d(ss) = B1;  db/dw = (-S1, C1).
    const ld dcc = -q.B1, dcs = q.A1, dss = q.B1;
    const ld db0 = -q.S1, db1 = q.C1;
    const ld mx0 = dcc * x0 + dcs * x1, mx1 = dcs * x0 + dss * x1;
    j.dp = x0 * db0 + x1 * db1 - (x0 * mx0 + x1 * mx1) / 2;
    // dx/dw solves M dx = db - dM x.
    const ld r0 = db0 - mx0, r1 = db1 - mx1;
    const ld dx0 = (ss * r0 - cs * r1) / det;
    const ld dx1 = (cc * r1 - cs * r0) / det;
    // d2M/dw2: d2(cc) = -2 A2, d2(cs) = -2 B2, d2(ss) = 2 A2;  d2b = (-C2, -S2).
    const ld d2cc = -2 * q.A2, d2cs = -2 * q.B2, d2ss = 2 * q.A2;
    const ld xm2x = d2cc * x0 * x0 + 2 * d2cs * x0 * x1 + d2ss * x1 * x1;
    j.d2p = (dx0 * r0 + dx1 * r1) + (x0 * -q.C2 + x1 * -q.S2) - xm2x / 2;
    return j;
}

}  // namespace

extern "C" {

// rust/src/ls.rs:30-74 at angular frequency w (theta = w t).
double or_lsd_tau_power(const double* t, const double* y, int64_t n, double w) {
    ld mc = 0, ms = 0, cc = 0, cs = 0;
    for (int64_t i = 0; i < n; ++i) {
        const ld th = static_cast<ld>(w) * static_cast<ld>(t[i]);
        const ld s = sinl(th), c = cosl(th);
        mc += y[i] * c;
        ms += y[i] * s;
        cc += c * c;
        cs += c * s;
    }
    const ld ss = static_cast<ld>(n) - cc;
    const ld tau = atan2l(2 * cs, cc - ss) / 2;
    const ld st = sinl(tau), ct = cosl(tau);
    ld v = ct * mc + st * ms;
    const ld num_l = v * v;
    v = ct * ms - st * mc;
    const ld num_r = v * v;
    const ld den_l = ct * ct * cc + 2 * ct * st * cs + st * st * ss;
    const ld den_r = ct * ct * ss - 2 * ct * st * cs + st * st * cc;
    if (den_l == 0 || den_r == 0) return 0.0;
    return static_cast<double>((num_l / den_l + num_r / den_r) / 2);
}

// out = {P, dP/dw, d2P/dw2}; dpdy[i] = dP/dy_i (may be NULL).
void or_lsd_jet(const double* t, const double* y, int64_t n, double w, double* out,
                double* dpdy) {
    const JetLd j = jet(sums(t, y, n, w), static_cast<ld>(n));
    out[0] = static_cast<double>(j.p);
    out[1] = static_cast<double>(j.dp);
    out[2] = static_cast<double>(j.d2p);
    if (dpdy)
        for (int64_t i = 0; i < n; ++i) {
            const ld th = static_cast<ld>(w) * static_cast<ld>(t[i]);
            dpdy[i] = static_cast<double>(j.x0 * cosl(th) + j.x1 * sinl(th));
        }
}

// Safeguarded Newton on dP/dw in [lo, hi] from w0 (the device algorithm).
// Returns status 0 converged, 1 no bracketed maximum, 2 not converged.
int or_lsd_refine(const double* t, const double* y, int64_t n, double lo, double hi, double w0,
                  double rtol, int max_iter, double* w_out, double* p_out) {
    ld a = lo, b = hi;
    if (a > b) {
        const ld tmp = a;
        a = b;
        b = tmp;
    }
    const ld nn = static_cast<ld>(n);
    const JetLd ja = jet(sums(t, y, n, a), nn), jb = jet(sums(t, y, n, b), nn);
    if (!(ja.dp > 0 && jb.dp < 0)) {
        *w_out = w0;
        *p_out = static_cast<double>(jet(sums(t, y, n, w0), nn).p);
        return 1;
    }
    ld w = w0;
    if (!(w > a && w < b)) w = (a + b) / 2;
    int st = 2;
    for (int it = 0; it < max_iter; ++it) {
        const JetLd j = jet(sums(t, y, n, w), nn);
        ld wn = w;
        if (j.dp != 0) {
            if (j.dp > 0) a = w;
            else b = w;
            wn = w - j.dp / j.d2p;
            if (!(j.d2p < 0) || !(wn > a && wn < b)) wn = (a + b) / 2;
        }
        const ld tol = static_cast<ld>(rtol) * fabsl(w);
        const bool conv = fabsl(wn - w) <= tol || (b - a) <= tol;
        w = wn;
        if (conv) {
            st = 0;
            break;
        }
    }
    *w_out = static_cast<double>(w);
    *p_out = static_cast<double>(jet(sums(t, y, n, w), nn).p);
    return st;
}

}  // extern "C"
