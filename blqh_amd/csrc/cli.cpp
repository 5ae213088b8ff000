// cli.cpp -- the blqh command line, a drop-in for the reference programs.
//
// Spinless (default build, `blqh`): flags and defaults of
// BLQH_without_spin/init.cpp:21-118, report of BLQH_without_spin/main.cpp:11-71.
// Spinful (-DBLQH_SPINFUL, `blqh_spin`): BLQH_with_spin/init.cpp:22-159 and
// BLQH_with_spin/main.cpp:11-99.
//
// The ED block runs the dense cross-check on the GPU; sectors above 16384
// states print "skipped" instead (the reference's dense block needs N^2 memory).
// Lanczos stops when the lowest Ritz value moves by less than BLQH_LANCZOS_TOL
// (default 1e-10) or after -m steps; BLQH_LANCZOS_TOL=0 runs exactly -m steps.
// -s sets the seed in both programs (the reference's spinless build ignores
// it); the default seed is clock-based as in the reference.
#include <unistd.h>

#include <algorithm>
#include <bitset>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <vector>

#include "blqh.h"

namespace {

void die_api(int rc, const char* what) {
    std::cerr << "blqh: " << what << " failed (" << rc << "): " << blqh_last_error() << std::endl;
    std::exit(1);
}

struct Run {
    int nLL, nphi, nel, nel_up, nel_su, J, kx, lambda, nthread;
    double d, dsas, dv, dz, gamma, theta;
    unsigned long seed;
};

#ifndef BLQH_SPINFUL
void usage(const char* target) {
    std::cout << "Usage: " << target << " [Options]\n";
    std::cout << "Options:\n";
    std::cout << "  -l                       nLL\n";
    std::cout << "  -n                       nphi\n";
    std::cout << "  -e                       Total No. of electrons in two layers\n";
    std::cout << "  -u                       No. of electrons in upper-layer\n";
    std::cout << "  -k                       kx in upper-layer\n";
    std::cout << "  -S                       Delta_SAS: tunnelling amplitude\n";
    std::cout << "  -v                       Delta_V: bias voltage\n";
    std::cout << "  -j                       total J in upper-layer or down-layer\n";
    std::cout << "  -g                       gamma=lx/ly  aspect ratio\n";
    std::cout << "  -d                       interlayer distance\n";
    std::cout << "  -m                       Lambda\n";
    std::cout << "  -t                       nthread\n";
    std::cout << "  -s                       seed\n";
    std::cout << "Default: (l,n,u,d,lambda) = (4,4,2,1,200)\n";
}

Run parse(int argc, char** argv) {
    Run r{};
    r.nLL = 0; r.nphi = 8; r.nel = 8; r.nel_up = -1; r.nel_su = -1; r.gamma = 1.0; r.nthread = 32;
    r.J = -1; r.kx = -1; r.dsas = 0; r.dv = 0; r.dz = 0; r.d = 100.0; r.lambda = 200; r.theta = 0;
    r.seed = (unsigned long)std::chrono::high_resolution_clock::now().time_since_epoch().count();
    int ch, err = 0;
    while ((ch = getopt(argc, argv, "L:e:n:u:d:j:k:g:m:t:s:v:h:S:")) != -1) {
        switch (ch) {
            case 'L': r.nLL = atoi(optarg); break;
            case 'n': r.nphi = atoi(optarg); break;
            case 'e': r.nel = atoi(optarg); break;
            case 'u': r.nel_up = atoi(optarg); break;
            case 'k': r.kx = atoi(optarg); break;
            case 'j': r.J = atoi(optarg); break;
            case 'g': r.gamma = atof(optarg); break;
            case 'd': r.d = atof(optarg); break;
            case 'S': r.dsas = atof(optarg); break;
            case 'v': r.dv = atof(optarg); break;
            case 'm': r.lambda = atoi(optarg); break;
            case 't': r.nthread = atoi(optarg); break;
            case 's': r.seed = (unsigned long)atoi(optarg); break;
            default: err++; break;
        }
    }
    try {
        if (r.nel < 0) throw std::logic_error("-n: positive value required !");
        if (r.nel_up > r.nel) throw std::logic_error("-u: nel_up < nel !");
    } catch (std::logic_error& e) {
        std::cout << e.what() << std::endl;
        usage(argv[0]);
        std::exit(2);
    }
    if (err) {
        usage(argv[0]);
        std::exit(0);
    }
    return r;
}
#else
void usage(const char* target) {
    std::cout << "Usage: " << target << " [Options]\n";
    std::cout << "Options:\n";
    std::cout << "  -L                       nLL\n";
    std::cout << "  -n                       nphi\n";
    std::cout << "  -e                       No. of electrons\n";
    std::cout << "  -U                       No. of upper-layer electrons\n";
    std::cout << "  -u                       No. of spin-up electrons\n";
    std::cout << "  -S                       Delta_SAS: tunnelling amplitude\n";
    std::cout << "  -V                       Delta_V: bias voltage\n";
    std::cout << "  -Z                       Delta_Z: Zeemann energy\n";
    std::cout << "  -J                       total J (K_y)\n";
    std::cout << "  -K                       Kx \n";
    std::cout << "  -g                       gamma=lx/ly  aspect ratio\n";
    std::cout << "  -d                       interlayer distance\n";
    std::cout << "  -m                       Lambda\n";
    std::cout << "  -s                       seed\n";
    std::cout << "  -t                       No. threads\n";
}

Run parse(int argc, char** argv) {
    Run r{};
    r.nLL = 0; r.nphi = 4; r.nel = 8; r.nel_up = -1; r.nel_su = -1; r.gamma = 1.0; r.theta = 0; r.nthread = 8;
    r.J = -1; r.kx = -1; r.dsas = 0; r.dz = 0; r.dv = 0; r.d = 1.0; r.lambda = 200;
    r.seed = (unsigned long)std::chrono::high_resolution_clock::now().time_since_epoch().count();
    int ch, err = 0;
    while ((ch = getopt(argc, argv, "L:e:n:U:u:d:J:K:g:Z:m:t:S:s:V:h:")) != -1) {
        switch (ch) {
            case 'L': r.nLL = atoi(optarg); break;
            case 'n': r.nphi = atoi(optarg); break;
            case 'e': r.nel = atoi(optarg); break;
            case 'U': r.nel_up = atoi(optarg); break;
            case 'u': r.nel_su = atoi(optarg); break;
            case 'K': r.kx = atoi(optarg); break;
            case 'J': r.J = atoi(optarg); break;
            case 'g': r.gamma = atof(optarg); break;
            case 'd': r.d = atof(optarg); break;
            case 'S': r.dsas = atof(optarg); break;
            case 's': r.seed = (unsigned long)atof(optarg); break;
            case 'V': r.dv = atof(optarg); break;
            case 'Z': r.dz = atof(optarg); break;
            case 'm': r.lambda = atoi(optarg); break;
            case 't': r.nthread = atoi(optarg); break;
            default: err++; break;
        }
    }
    auto bail = [&](const char* msg) {
        std::cout << msg << std::endl;
        usage(argv[0]);
        std::exit(2);
    };
    if (r.nphi < 0) bail("-n: positive value required !");
    if (r.nel <= 0 || r.nel >= 4 * r.nphi) bail("-e: Total No. of electrons should be in (0,4*nphi)");
    if (r.nel_up < 0 || r.nel_up >= r.nel) bail("-U: No. of upper-layer electrons should be in [0,nel_up] !");
    if (r.nel_su < 0 || r.nel_su >= r.nel) bail("-u: No. of spin-up electrons should be in [0,nel_up] !");
    if (r.nphi > 21 && r.nel > 2 * r.nphi / 3 && r.nel < r.nphi * 4 / 3) bail("Hilbert space too large !");
    if (err) {
        usage(argv[0]);
        std::exit(0);
    }
    return r;
}
#endif

}  // namespace

int main(int argc, char** argv) {
    Run r = parse(argc, argv);
#ifdef BLQH_SPINFUL
    const int S = 4;
#else
    const int S = 2;
#endif
    const double ly = std::sqrt(r.nphi * 2.0 * M_PI / r.gamma);
    const double lx = ly * r.gamma;
    // The reference runs exactly lambda Lanczos steps (no convergence test),
    // so by default the drop-in does too (its printed amplitudes then carry
    // the reference's digits); BLQH_LANCZOS_TOL > 0 stops at |dE| < tol instead.
    double tol = 0.0;
    if (const char* e = std::getenv("BLQH_LANCZOS_TOL")) tol = std::atof(e);

    blqh_sector sec{S, r.nphi, r.nel, r.nel_up, S == 4 ? r.nel_su : -1, r.J, r.kx};
    blqh_params par{lx, ly, r.nLL, r.d, r.dsas, r.dv, r.dz, 0.0, 0.0, 0.0};
    int rc;
    blqh_basis* B = nullptr;
    std::cout << "----------- ED results --------------" << std::endl;
    if ((rc = blqh_basis_create(&sec, &B))) die_api(rc, "basis");
    const int64_t n = blqh_basis_size(B);
    std::cout << "nHilbert: =" << n << std::endl;
    blqh_hamil* H = nullptr;
    if ((rc = blqh_hamil_build(B, &par, &H))) die_api(rc, "hamiltonian build");
    if (n <= 16384) {
        double e0 = 0, ep = 0;
        if ((rc = blqh_dense_ground_state(H, &e0, &ep, nullptr))) die_api(rc, "dense diagonalisation");
        std::cout << "E_gs:= " << std::setprecision(S == 2 ? 6 : 10) << ep / r.nel << std::endl;
    } else {
        std::cout << "E_gs:= skipped (nHilbert > 16384)" << std::endl;
    }
    std::cout << "----------- Lanczos results ---------" << std::endl;
    blqh_lanczos_opts lo{r.lambda, (uint32_t)r.seed, tol, 1, -1, nullptr, tol > 0 ? 1e-7 : 0.0};
    blqh_lanczos_result res{};
    std::vector<double> psi(2 * n);
    if ((rc = blqh_lanczos(H, &lo, &res, psi.data(), nullptr, nullptr))) die_api(rc, "lanczos");
    std::cout << "E_gs:= " << std::setprecision(10) << res.E_rq / r.nel << std::endl;
    std::vector<uint64_t> id(n);
    blqh_basis_export(B, id.data(), nullptr);
    std::vector<std::pair<long, double>> wf;
    wf.reserve(n);
    for (long i = 0; i < n; ++i) wf.emplace_back(i, std::abs(std::complex<double>(psi[2 * i], psi[2 * i + 1])));
    std::sort(wf.begin(), wf.end(), [](const std::pair<long, double>& a, const std::pair<long, double>& b) {
        return a.second > b.second;
    });
    const int nphi = r.nphi;
#ifndef BLQH_SPINFUL
    std::cout << "# ground state wave function" << std::endl;
    int count = 0;
    for (const auto& p : wf) {
        ++count;
        std::cout << std::setw(4) << p.first << " :   |";
        const unsigned long long u = id[p.first];
        for (int k = 0; k < nphi; ++k)
            if ((u >> k) & 1ULL) std::cout << std::setw(3) << k; else std::cout << std::setw(3) << "_";
        std::cout << ")|";
        for (int k = nphi; k < 2 * nphi; ++k)
            if ((u >> k) & 1ULL) std::cout << std::setw(3) << k - nphi; else std::cout << std::setw(3) << "_";
        std::cout << ")   " << std::bitset<24>(id[p.first]).to_string() << "  " << p.second << std::endl;
        if (count > 10 || p.second < 1e-3) break;
    }
#else
    int count = 0;
    for (const auto& p : wf) {
        ++count;
        std::cout << std::setw(10) << "(" << p.first << "): " << std::bitset<32>(id[p.first]).to_string() << "  "
                  << p.second << std::endl;
        std::cout << std::setw(10) << "upper layer:   |";
        const unsigned long long u = id[p.first];
        for (int k = 0; k < nphi; ++k)
            if ((u >> k) & 1ULL) std::cout << std::setw(3) << k; else std::cout << std::setw(3) << "_";
        std::cout << ")|";
        for (int k = nphi; k < 2 * nphi; ++k)
            if ((u >> k) & 1ULL) std::cout << std::setw(3) << k - nphi; else std::cout << std::setw(3) << "_";
        std::cout << ")|" << std::endl;
        std::cout << std::setw(10) << " down layer:   |";
        for (int k = 2 * nphi; k < 3 * nphi; ++k)
            if ((u >> k) & 1ULL) std::cout << std::setw(3) << k; else std::cout << std::setw(3) << "_";
        std::cout << ")|";
        for (int k = 3 * nphi; k < 4 * nphi; ++k)
            if ((u >> k) & 1ULL) std::cout << std::setw(3) << k; else std::cout << std::setw(3) << "_";
        std::cout << ")   " << std::endl;
        if (count > n || p.second < 1e-5 || count > 10) break;
    }
#endif
    blqh_hamil_destroy(H);
    blqh_basis_destroy(B);
    return 0;
}
