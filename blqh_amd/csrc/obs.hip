// obs.hip -- ground-state observables.
//
// Reference: lhamil::occupatation_number, pseudospin_Sz, pseudospin_Sx
// (BLQH_without_spin/lanczos_hamiltonian.cpp:658-735) and the spinful
// occupatation_number, pseudospin_Sz, Sz, upper_Sz, down_Sz, pseudospin_Sx
// (BLQH_with_spin/lanczos_hamiltonian.cpp:1217-1500), evaluated on a state
// psi given in the basis of a sector:
//   occupation(a, j) = sum_i |psi_i|^2 / D_i * #{q < D_i : bit of orbital j of
//                      layer a set in T^q id_i}        (a = 0: bits j, a = 1: bits j + nphi)
//   pseudospin Sz    = (2 <N_upper> - Ne) / (2 Ne)     (spinful: upper = species 0, 1)
//   Sz (spinful)     = (2 <N_spin-up> - Ne) / (2 Ne)   (species 0, 2)
//   upper / lower Sz = < (2 n_0 - n_U) / (2 n_U + 1e-8) >, < (2 n_2 - n_L) / (2 n_L + 1e-8) >
//   pseudospin Sx    = | <psi| T |psi> | / Ne
//   spin flip        = | <psi| T_sf |psi> | / Ne   (spinful spinflip_tunneling, :1519-1592)
// T is the tunnelling operator of the Hamiltonian with coefficient +1/2 per
// orbital pair (spinless: n <-> n + nphi; spinful: species pairs (0,2), (1,3),
// (0,3), (1,2), each visited from both ends with 1/4); T_sf keeps the
// spin-flip pairs (0,3) and (1,2), each visited once with 1/2.  Both are evaluated row by row
// with the Hamiltonian build's hop term (translation images, representative
// lookup, fermion signs, FT phases).  Sums run per thread in row order and
// per block in a fixed order; the host adds the block partials in order.
#include <cmath>
#include <cstring>
#include <vector>

#include "engine.hpp"

namespace {

using namespace blqh;

constexpr int kT = 256;

__device__ __forceinline__ double block_sum(double v, double* sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0;
    for (int w = 0; w < kT / 64; ++w) t += sh[w];
    __syncthreads();
    return t;
}

// <psi| T |psi>: partials (re, im) per block
__global__ void __launch_bounds__(kT) k_obs_hops(RowCtx c, int64_t n, const uint64_t* __restrict__ id,
                                                 const HopTerm* __restrict__ hops, int nh,
                                                 const double2* __restrict__ psi, double2* __restrict__ part) {
    __shared__ double sh[kT / 64];
    double ar = 0.0, ai = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
        const int Dl = c.K < 0 ? 1 : c.period[i];
        const double2 pl = psi[i];
        for (int ql = 0; ql < Dl; ++ql) {
            int signl = 1;
            const uint64_t L = ql ? translate(c.g, id[i], ql, signl) : id[i];
            for (int h = 0; h < nh; ++h) {
                int j;
                double re, im;
                if (!hop_term(c, L, signl, Dl, ql, hops[h], j, re, im)) continue;
                const double2 pr = psi[j];
                // conj(psi_i) * (re + i im) * psi_j
                const double tr = re * pr.x - im * pr.y, ti = re * pr.y + im * pr.x;
                ar += pl.x * tr + pl.y * ti;
                ai += pl.x * ti - pl.y * tr;
            }
        }
    }
    const double sr = block_sum(ar, sh), si = block_sum(ai, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = make_double2(sr, si);
}

// occupation numbers: blockIdx.y = mask index (orbital j of layer a = y / nphi)
__global__ void __launch_bounds__(kT) k_obs_occ(Geo g, int K, int64_t n, const uint64_t* __restrict__ id,
                                                const int16_t* __restrict__ period,
                                                const double2* __restrict__ psi, double* __restrict__ part) {
    __shared__ double sh[kT / 64];
    const int y = blockIdx.y;
    const uint64_t mask = 1ULL << y;  // orbital y % nphi of layer y / nphi
    double acc = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
        const int D = K < 0 ? 1 : period[i];
        int cnt = 0;
        for (int q = 0; q < D; ++q) {
            int s;
            const uint64_t w = q ? translate(g, id[i], q, s) : id[i];
            cnt += (w & mask) == mask;
        }
        const double2 z = psi[i];
        acc += (z.x * z.x + z.y * z.y) * cnt / D;
    }
    const double t = block_sum(acc, sh);
    if (threadIdx.x == 0) part[(int64_t)y * gridDim.x + blockIdx.x] = t;
}

// <N_upper>, <N_spin-up>, <upper Sz>, <lower Sz> partials (4 per block)
__global__ void __launch_bounds__(kT) k_obs_sz(Geo g, int64_t n, const uint64_t* __restrict__ id,
                                               const double2* __restrict__ psi, double* __restrict__ part) {
    __shared__ double sh[kT / 64];
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    const uint64_t m = lowbits(g.nphi);
    for (int64_t i = blockIdx.x * (int64_t)kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
        const uint64_t w = id[i];
        const double2 z = psi[i];
        const double p = z.x * z.x + z.y * z.y;
        const int n0 = popc(w & m), n1 = popc((w >> g.nphi) & m);
        if (g.S == 2) {
            a0 += n0 * p;
        } else {
            const int n2 = popc((w >> 2 * g.nphi) & m), n3 = popc((w >> 3 * g.nphi) & m);
            a0 += (n0 + n1) * p;
            a1 += (n0 + n2) * p;
            const double nu = n0 + n1, nl = n2 + n3;
            a2 += (2.0 * n0 - nu) / (2.0 * nu + 1e-8) * p;
            a3 += (2.0 * n2 - nl) / (2.0 * nl + 1e-8) * p;
        }
    }
    const double s0 = block_sum(a0, sh), s1 = block_sum(a1, sh), s2 = block_sum(a2, sh), s3 = block_sum(a3, sh);
    if (threadIdx.x == 0) {
        part[4 * blockIdx.x] = s0;
        part[4 * blockIdx.x + 1] = s1;
        part[4 * blockIdx.x + 2] = s2;
        part[4 * blockIdx.x + 3] = s3;
    }
}

// tunnelling operator T of the observables (coefficient 1/2 per pair)
std::vector<HopTerm> sx_hops(const Geo& g) {
    std::vector<HopTerm> h;
    const int f = g.nphi;
    auto add = [&](int a, int b, double coef) {
        h.push_back({a, b, a < b ? a : b, a < b ? b : a, coef, 1.0, 0.0, 0});
    };
    if (g.S == 2) {
        for (int n = 0; n < f; ++n) add(n, n + f, 0.5);
    } else {
        for (int n = 0; n < f; ++n) add(n, n + 2 * f, 0.25);
        for (int n = f; n < 2 * f; ++n) add(n, n + 2 * f, 0.25);
        for (int n = 2 * f; n < 3 * f; ++n) add(n, n - 2 * f, 0.25);
        for (int n = 3 * f; n < 4 * f; ++n) add(n, n - 2 * f, 0.25);
        for (int n = 0; n < f; ++n) add(n, n + 3 * f, 0.25);
        for (int n = 3 * f; n < 4 * f; ++n) add(n, n - 3 * f, 0.25);
        for (int n = f; n < 2 * f; ++n) add(n, n + f, 0.25);
        for (int n = 2 * f; n < 3 * f; ++n) add(n, n - f, 0.25);
    }
    return h;
}

// spin-flip part of the interlayer tunnelling (spinful only; coefficient 1/2 per pair)
std::vector<HopTerm> sf_hops(const Geo& g) {
    std::vector<HopTerm> h;
    const int f = g.nphi;
    auto add = [&](int a, int b) { h.push_back({a, b, a < b ? a : b, a < b ? b : a, 0.5, 1.0, 0.0, 0}); };
    for (int n = 0; n < f; ++n) add(n, n + 3 * f);      // upper-up <-> lower-down
    for (int n = f; n < 2 * f; ++n) add(n, n + f);      // upper-down <-> lower-up
    return h;
}

}  // namespace

void blqh_observables_impl(const blqh_basis* B, const double* psi_host, blqh_observables& out) {
    hipStream_t st = stream();
    const int64_t n = B->n;
    const Geo& g = B->geo;
    BLQH_REQUIRE(g.nphi <= 64, BLQH_EINVAL, "too many orbitals");
    BLQH_REQUIRE(g.S >= 2, BLQH_EINVAL, "the pseudospin and spin observables need a bilayer basis");
    DevBuf<double2> psi(std::max<int64_t>(n, 1));
    BLQH_HIP(hipMemcpyAsync(psi.get(), psi_host, n * sizeof(double2), hipMemcpyHostToDevice, st));
    // FT phases of the sector
    const int f = g.nphi, K = B->sec.K;
    std::vector<double2> ft(2 * (size_t)f - 1);  // indexed by ql - qr + f - 1
    for (int dq = -f + 1; dq < f; ++dq)
        ft[dq + f - 1] = make_double2(std::cos(2.0 * M_PI * dq * K / f), std::sin(2.0 * M_PI * dq * K / f));
    DevBuf<double2> dft(ft.size());
    BLQH_HIP(hipMemcpyAsync(dft.get(), ft.data(), ft.size() * sizeof(double2), hipMemcpyHostToDevice, st));
    std::vector<HopTerm> hops = sx_hops(g);
    const int nsx = (int)hops.size();
    if (g.S == 4) {
        const std::vector<HopTerm> sf = sf_hops(g);
        hops.insert(hops.end(), sf.begin(), sf.end());
    }
    const int nsf = (int)hops.size() - nsx;
    DevBuf<HopTerm> dh(hops.size());
    BLQH_HIP(hipMemcpyAsync(dh.get(), hops.data(), hops.size() * sizeof(HopTerm), hipMemcpyHostToDevice, st));
    RowCtx c = make_row_ctx(B);
    c.FT = dft.get();
    const int G = (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, kT), 2048));
    DevBuf<double2> ph(2 * (size_t)G);
    DevBuf<double> po((size_t)G * 2 * f), pz((size_t)G * 4);
    hipLaunchKernelGGL(k_obs_hops, dim3(G), dim3(kT), 0, st, c, n, B->id.get(), dh.get(), nsx, psi.get(), ph.get());
    if (nsf)
        hipLaunchKernelGGL(k_obs_hops, dim3(G), dim3(kT), 0, st, c, n, B->id.get(), dh.get() + nsx, nsf, psi.get(),
                           ph.get() + G);
    hipLaunchKernelGGL(k_obs_occ, dim3(G, 2 * f), dim3(kT), 0, st, g, K, n, B->id.get(), B->period.get(), psi.get(),
                       po.get());
    hipLaunchKernelGGL(k_obs_sz, dim3(G), dim3(kT), 0, st, g, n, B->id.get(), psi.get(), pz.get());
    BLQH_HIP(hipGetLastError());
    std::vector<double2> hh(2 * (size_t)G);
    std::vector<double> ho((size_t)G * 2 * f), hz((size_t)G * 4);
    BLQH_HIP(hipMemcpyAsync(hh.data(), ph.get(), (nsf ? 2 : 1) * G * sizeof(double2), hipMemcpyDeviceToHost, st));
    BLQH_HIP(hipMemcpyAsync(ho.data(), po.get(), ho.size() * 8, hipMemcpyDeviceToHost, st));
    BLQH_HIP(hipMemcpyAsync(hz.data(), pz.get(), hz.size() * 8, hipMemcpyDeviceToHost, st));
    BLQH_HIP(hipStreamSynchronize(st));
    double tr = 0, ti = 0, fr = 0, fi = 0, z[4] = {0, 0, 0, 0};
    for (int b = 0; b < G; ++b) {
        tr += hh[b].x;
        ti += hh[b].y;
        if (nsf) {
            fr += hh[G + b].x;
            fi += hh[G + b].y;
        }
        for (int k = 0; k < 4; ++k) z[k] += hz[4 * b + k];
    }
    std::memset(&out, 0, sizeof(out));
    const double ne = g.nel;
    out.pseudospin_sx = std::hypot(tr, ti) / ne;
    out.pseudospin_sz = (2.0 * z[0] - ne) / (2.0 * ne);
    if (g.S == 4) {
        out.sz = (2.0 * z[1] - ne) / (2.0 * ne);
        out.upper_sz = z[2];
        out.lower_sz = z[3];
        out.spinflip = std::hypot(fr, fi) / ne;
    }
    for (int y = 0; y < 2 * f; ++y) {
        double s = 0;
        for (int b = 0; b < G; ++b) s += ho[(size_t)y * G + b];
        out.occupation[y / f][y % f] = s;
    }
}
