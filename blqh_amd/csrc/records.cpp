// records.cpp -- Lanczos spectra and the Lanczos record file.
//
// Spectrum: every eigenpair of the Lanczos tridiagonal T (the reference's
// diag(), BLQH_without_spin/lanczos_hamiltonian.cpp:581-606, which calls
// LAPACK dsyevd on the dense T), here by the implicit-shift QL algorithm for
// symmetric tridiagonal matrices (Bowdler, Martin, Reinsch and Wilkinson,
// Numer. Math. 11 (1968) 293, "tql2").  The low-lying spectrum keeps the
// first eigenvalue and every later one that exceeds its predecessor by more
// than a gap (the ghost filter of the reference's spectrum sweeps,
// BLQH_without_spin/main.cpp:100-110: Lanczos without reorthogonalisation
// repeats converged eigenvalues).
//
// Record: the text format of lhamil::save_to_file / read_from_file
// (BLQH_without_spin/lanczos_hamiltonian.cpp:738-819):
//   nHilbert:=<setw(10) " "><n>        lambda:= ...   seed:= ...
//   eigenvalues:=<setw(10) " [ "> e0 , e1 , ... ,  ]     (8 significant digits)
//   norm:= [ beta_0 .. beta_m ]  overlap:= [ alpha ]  psi_0:= [...]  psi_n0:= [...]
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "engine.hpp"

namespace blqh {

// All eigenpairs of the symmetric tridiagonal matrix with diagonal d[m] and
// off-diagonal e[m-1]; on return d holds the eigenvalues ascending and z
// (m x m, column k = eigenvector k, z[i + m k]) the orthonormal eigenvectors.
void tridiag_eigen(int m, std::vector<double>& d, std::vector<double> e, std::vector<double>& z) {
    z.assign((size_t)m * m, 0.0);
    for (int i = 0; i < m; ++i) z[i + (size_t)m * i] = 1.0;
    e.resize(m);
    e[m - 1] = 0.0;  // e[i] couples rows i and i+1
    for (int l = 0; l < m; ++l) {
        int iter = 0;
        for (;;) {
            int mm = l;
            for (; mm + 1 < m; ++mm) {
                const double dd = std::fabs(d[mm]) + std::fabs(d[mm + 1]);
                if (std::fabs(e[mm]) <= 1e-15 * dd) break;
            }
            if (mm == l) break;
            BLQH_REQUIRE(++iter <= 60, BLQH_EINTERNAL, "tridiagonal QL did not converge");
            // Wilkinson-like shift from the leading 2x2 block
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[mm] - d[l] + e[l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
            double s = 1.0, c = 1.0, p = 0.0;
            int i = mm - 1;
            for (; i >= l; --i) {
                double f = s * e[i];
                const double b = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {  // deflation inside the sweep
                    d[i + 1] -= p;
                    e[mm] = 0.0;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                for (int k = 0; k < m; ++k) {  // rotate eigenvector columns i, i+1
                    f = z[k + (size_t)m * (i + 1)];
                    z[k + (size_t)m * (i + 1)] = s * z[k + (size_t)m * i] + c * f;
                    z[k + (size_t)m * i] = c * z[k + (size_t)m * i] - s * f;
                }
            }
            if (r == 0.0 && i >= l) continue;
            d[l] -= p;
            e[l] = g;
            e[mm] = 0.0;
        }
    }
    // ascending order (selection sort, columns follow)
    for (int i = 0; i + 1 < m; ++i) {
        int k = i;
        for (int j = i + 1; j < m; ++j)
            if (d[j] < d[k]) k = j;
        if (k != i) {
            std::swap(d[i], d[k]);
            for (int r = 0; r < m; ++r) std::swap(z[r + (size_t)m * i], z[r + (size_t)m * k]);
        }
    }
}

}  // namespace blqh

namespace {

std::string fmt8(double v) {
    char b[64];
    std::snprintf(b, sizeof b, "%.8g", v);  // ostream defaultfloat with precision 8
    return b;
}

void put_list(std::ostream& o, const char* name, const double* v, int count) {
    char pad[16];
    std::snprintf(pad, sizeof pad, "%10s", " [ ");  // setw(10) << " [ "
    o << name << pad;
    for (int i = 0; i < count; ++i) o << fmt8(v[i]) << " , ";
    o << " ] " << "\n";
}

}  // namespace

void blqh_tridiag_spectrum_impl(int m, const double* alpha, const double* offdiag, double* eig, double* first,
                                double* vec0) {
    BLQH_REQUIRE(m >= 1, BLQH_EINVAL, "empty tridiagonal");
    std::vector<double> d(alpha, alpha + m), e(m > 1 ? m - 1 : 1, 0.0), z;
    if (m > 1) std::memcpy(e.data(), offdiag, (m - 1) * sizeof(double));
    blqh::tridiag_eigen(m, d, e, z);
    // sign convention of the ground-state vector: first component non-negative
    const double sg = z[0] < 0 ? -1.0 : 1.0;
    for (int k = 0; k < m; ++k) {
        if (eig) eig[k] = d[k];
        if (first) first[k] = z[(size_t)m * k];      // psi_n0: first component of eigenvector k
        if (vec0) vec0[k] = sg * z[k];               // psi_0: eigenvector 0
    }
}

void blqh_tridiag_vectors_impl(int m, const double* alpha, const double* offdiag, int k0, int count, double* vecs) {
    BLQH_REQUIRE(m >= 1 && k0 >= 0 && count >= 1 && k0 + count <= m, BLQH_EINVAL, "eigenvector index out of range");
    std::vector<double> d(alpha, alpha + m), e(m > 1 ? m - 1 : 1, 0.0), z;
    if (m > 1) std::memcpy(e.data(), offdiag, (m - 1) * sizeof(double));
    blqh::tridiag_eigen(m, d, e, z);
    for (int k = k0; k < k0 + count; ++k) {
        const double* col = z.data() + (size_t)m * k;
        const double sg = col[0] < 0 ? -1.0 : 1.0;
        for (int i = 0; i < m; ++i) vecs[(size_t)(k - k0) * m + i] = sg * col[i];
    }
}

int blqh_spectrum_distinct_impl(int m, const double* eig, int kmax, double gap, double* out) {
    // eigenvalues[0], then eigenvalues[i] with eigenvalues[i] - eigenvalues[i-1] > gap
    int count = 0;
    if (m < 1 || kmax < 1) return 0;
    out[count++] = eig[0];
    for (int i = 1; i < m && count < kmax; ++i)
        if (eig[i] - eig[i - 1] > gap) out[count++] = eig[i];
    return count;
}

void blqh_record_save_impl(const char* path, int64_t nhilbert, int m, uint32_t seed, const double* eig,
                           const double* beta, const double* alpha, const double* psi0, const double* psin0) {
    std::ofstream o(path);
    BLQH_REQUIRE(o.good(), BLQH_EINVAL, std::string("cannot open ") + path);
    char pad[16];
    std::snprintf(pad, sizeof pad, "%10s", " ");
    o << "nHilbert:=" << pad << nhilbert << "\n";
    o << "lambda:=" << pad << m << "\n";
    o << "seed:=" << pad << seed << "\n";
    put_list(o, "eigenvalues:=", eig, m);
    put_list(o, "norm:=", beta, m + 1);
    put_list(o, "overlap:=", alpha, m);
    put_list(o, "psi_0:=", psi0, m);
    put_list(o, "psi_n0:=", psin0, m);
    BLQH_REQUIRE(o.good(), BLQH_EINTERNAL, std::string("write failed: ") + path);
}

// Reads a record the way read_from_file does (whitespace-separated tokens).
void blqh_record_load_impl(const char* path, int cap, int64_t* nhilbert, int* m, uint32_t* seed, double* eig,
                           double* beta, double* alpha, double* psi0, double* psin0) {
    std::ifstream in(path);
    BLQH_REQUIRE(in.good(), BLQH_EINVAL, std::string("cannot open ") + path);
    std::string tok;
    int64_t n = 0;
    long long lam = 0;
    unsigned long long sd = 0;
    in >> tok >> n >> tok >> lam >> tok >> sd;
    BLQH_REQUIRE(in.good() && lam >= 1, BLQH_EINVAL, "malformed record header");
    BLQH_REQUIRE(lam <= cap, BLQH_ERANGE, "record longer than the caller's arrays");
    auto list = [&](double* out, long long count) {
        in >> tok >> tok;  // name, "["
        for (long long i = 0; i < count; ++i) {
            double v;
            in >> v >> tok;
            BLQH_REQUIRE(!in.fail(), BLQH_EINVAL, "malformed record list");
            if (out) out[i] = v;
        }
        in >> tok;  // "]"
    };
    list(eig, lam);
    list(beta, lam + 1);
    list(alpha, lam);
    list(psi0, lam);
    list(psin0, lam);
    *nhilbert = n;
    *m = (int)lam;
    *seed = (uint32_t)sd;
}
