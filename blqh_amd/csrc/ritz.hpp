// ritz.hpp -- host routines for low Ritz vectors of a Lanczos tridiagonal,
// shared by the batched solver and its CPU tests.
#pragma once
#include <algorithm>
#include <cmath>
#include <vector>

namespace blqh {

// Number of eigenvalues below x of the tridiagonal (a, b) (Sturm sequence).
inline int sturm_below(int m, const double* a, const double* b, double x) {
    int c = 0;
    double q = 1.0;
    for (int i = 0; i < m; ++i) {
        q = a[i] - x - (i ? b[i - 1] * b[i - 1] / q : 0.0);
        if (q == 0.0) q = -1e-300;
        if (q < 0) ++c;
    }
    return c;
}

// Eigenvector k (ascending) of the m x m tridiagonal (a, b): the eigenvalue
// by bisection on Sturm counts, the vector by three steps of inverse
// iteration, each solve by Gaussian elimination with partial pivoting on the
// band (the steps of LAPACK's dgttrf / dgttrs).  x gets a unit vector.
inline void tridiag_vector_k(int m, const double* a, const double* b, int k, double* x) {
    double lo = a[0], hi = a[0];
    for (int i = 0; i < m; ++i) {
        const double r = (i ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < m ? std::fabs(b[i]) : 0.0);
        lo = std::min(lo, a[i] - r);
        hi = std::max(hi, a[i] + r);
    }
    for (int it = 0; it < 200; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        if (sturm_below(m, a, b, mid) >= k + 1) hi = mid; else lo = mid;
    }
    const double e = 0.5 * (lo + hi);
    const double sh = e - 1e-13 * std::max(1.0, std::fabs(e));
    std::vector<double> dl(m), dd(m), du(m), du2(m), y(m);
    std::vector<char> sw(m, 0);
    for (int i = 0; i < m; ++i) {
        dd[i] = a[i] - sh;
        du[i] = i + 1 < m ? b[i] : 0.0;
        dl[i] = i + 1 < m ? b[i] : 0.0;
        du2[i] = 0.0;
    }
    for (int i = 0; i + 1 < m; ++i) {
        if (std::fabs(dd[i]) >= std::fabs(dl[i])) {
            if (dd[i] == 0.0) dd[i] = 1e-300;
            const double l = dl[i] / dd[i];
            dl[i] = l;
            dd[i + 1] -= l * du[i];
        } else {  // rows i and i + 1 swap
            const double l = dd[i] / dl[i];
            dd[i] = dl[i];
            dl[i] = l;
            const double t = du[i];
            du[i] = dd[i + 1];
            dd[i + 1] = t - l * dd[i + 1];
            if (i + 2 < m) {
                du2[i] = du[i + 1];
                du[i + 1] = -l * du[i + 1];
            }
            sw[i] = 1;
        }
    }
    if (dd[m - 1] == 0.0) dd[m - 1] = 1e-300;
    for (int i = 0; i < m; ++i) x[i] = 1.0 / std::sqrt((double)m);
    for (int iter = 0; iter < 3; ++iter) {
        for (int i = 0; i < m; ++i) y[i] = x[i];
        for (int i = 0; i + 1 < m; ++i) {
            if (sw[i]) {
                const double t = y[i];
                y[i] = y[i + 1];
                y[i + 1] = t - dl[i] * y[i];
            } else {
                y[i + 1] -= dl[i] * y[i];
            }
        }
        for (int i = m - 1; i >= 0; --i) {
            double v = y[i];
            if (i + 1 < m) v -= du[i] * x[i + 1];
            if (i + 2 < m) v -= du2[i] * x[i + 2];
            x[i] = v / dd[i];
        }
        double s = 0;
        for (int i = 0; i < m; ++i) s += x[i] * x[i];
        s = 1.0 / std::sqrt(s);
        for (int i = 0; i < m; ++i) x[i] *= s;
    }
}

// Excited mix of a finished Lanczos run: the normalised sum of the second
// and third lowest Ritz vectors (those that exist), orthogonalised against
// the ground Ritz vector c0; zero when m = 1.  The start of a continuation
// solve adds it so that a state that was excited at the previous parameter
// point and is the ground state at the next one starts with a large weight.
inline void excited_mix(int m, const double* a, const double* b, const double* c0, double* out) {
    for (int l = 0; l < m; ++l) out[l] = 0.0;
    if (m < 2) return;
    std::vector<double> v(m);
    for (int k = 1; k <= std::min(2, m - 1); ++k) {
        tridiag_vector_k(m, a, b, k, v.data());
        for (int l = 0; l < m; ++l) out[l] += v[l];
    }
    double d = 0, s = 0;
    for (int l = 0; l < m; ++l) d += out[l] * c0[l];
    for (int l = 0; l < m; ++l) {
        out[l] -= d * c0[l];
        s += out[l] * out[l];
    }
    if (!(s > 0) || !std::isfinite(s)) {
        for (int l = 0; l < m; ++l) out[l] = 0.0;
        return;
    }
    s = 1.0 / std::sqrt(s);
    for (int l = 0; l < m; ++l) out[l] *= s;
}

}  // namespace blqh
