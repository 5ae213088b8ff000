// mlp.hip -- learned warm start for the Lanczos solver.
//
// No counterpart in the reference: it replaces the seeded random start vector
// of pass 1 (BLQH_without_spin/lanczos_hamiltonian.cpp:466-483, Vec::init_random
// matrix.cpp:54-75) by a network's guess of the ground state; the seeded start
// stays available and every solve still runs to the same stopping rule.
//
// Network (FP32): inputs x = (2 b_k - 1 for the occupation bits b_k of the
// basis word, k < nbits; d / 3; 10 Delta_SAS), two hidden layers of 64 tanh
// units, two outputs: o0 = log |psi_i| and o1 = logit of sign(psi_i) > 0.
// Parameters, flat: W1[nin][64], b1[64], W2[64][64], b2[64], W3[64][2], b3[2].
//
// Training data: ground states psi of (sector, d, Delta_SAS) sets computed by
// the engine, gauge-fixed to real (the phase of the largest component is
// removed).  A minibatch draws states with probability |psi_i|^2 within a set
// chosen uniformly, so the loss
//     L = mean_s [ (o0_s - log|psi_s|)^2 + softplus(o1_s) - sign01_s o1_s ]
// weighs configurations by their weight in the ground state.
// Step kernels: k_sample (draw the minibatch), k_fwd_bwd (a workgroup per tile
// of 64 states: forward, loss, back-propagation and the tile's gradient sums,
// all in LDS), k_grad_sum (per parameter, tiles summed in order:
// deterministic), k_adam.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "engine.hpp"

namespace {

using namespace blqh;

constexpr int kH = 64;      // hidden width
constexpr int kMaxIn = 66;  // 64 occupation bits + 2 parameters
constexpr int kT = 256;

struct Dims {
    int nbits, nin;
    int oW1, ob1, oW2, ob2, oW3, ob3, n;  // offsets into the flat parameter vector
};

Dims make_dims(int nbits) {
    Dims d;
    d.nbits = nbits;
    d.nin = nbits + 2;
    d.oW1 = 0;
    d.ob1 = d.oW1 + d.nin * kH;
    d.oW2 = d.ob1 + kH;
    d.ob2 = d.oW2 + kH * kH;
    d.oW3 = d.ob2 + kH;
    d.ob3 = d.oW3 + kH * 2;
    d.n = d.ob3 + 2;
    return d;
}

// one minibatch in HBM, struct-of-arrays
struct Samples {
    uint64_t* word;
    float* pd;     // d
    float* ps;     // Delta_SAS
    float* ya;     // log |psi|
    float* ys;     // 1 if psi > 0 else 0
    float* loss;   // [B]
};

__device__ __forceinline__ float feature(const Dims& D, uint64_t w, float pd, float ps, int k) {
    if (k < D.nbits) return ((w >> k) & 1ull) ? 1.0f : -1.0f;
    return k == D.nbits ? pd * (1.0f / 3.0f) : ps * 10.0f;
}

// A workgroup of kT threads evaluates a tile of kTile states.  LDS holds the
// parameters (W2 with a padded row stride so that a wavefront reading one
// column touches 64 different banks), the tile's inputs and activations.
constexpr int kTile = 64;
constexpr int kLd = kH + 1;  // padded row stride of the 64-wide LDS matrices
struct Lds {
    float W1[kMaxIn * kH], b1[kH], W2[kH * kLd], b2[kH], W3[kH * 2], b3[2];
    float X[kTile * (kMaxIn + 1)];  // inputs, row stride nin + 1
    float H1[kTile * kLd], H2[kTile * kLd], D2[kTile * kLd];
    float O[kTile * 2];
};

__device__ __forceinline__ void stage_params(const Dims& D, const float* __restrict__ P, Lds& L) {
    for (int i = threadIdx.x; i < D.nin * kH; i += kT) L.W1[i] = P[D.oW1 + i];
    for (int i = threadIdx.x; i < kH * kH; i += kT) L.W2[(i / kH) * kLd + i % kH] = P[D.oW2 + i];
    for (int i = threadIdx.x; i < kH; i += kT) {
        L.b1[i] = P[D.ob1 + i];
        L.b2[i] = P[D.ob2 + i];
    }
    for (int i = threadIdx.x; i < 2 * kH; i += kT) L.W3[i] = P[D.oW3 + i];
    if (threadIdx.x < 2) L.b3[threadIdx.x] = P[D.ob3 + threadIdx.x];
}

// Forward of a tile (nt <= kTile valid states, the rest zero inputs): L.H1,
// L.H2 (tanh activations), L.O (outputs).  Thread t computes unit t % 64 for
// states t / 64, t / 64 + 4, ...: a wavefront shares the state (broadcast
// reads of X / H) and spans the units (consecutive weight reads).
__device__ __forceinline__ void tile_forward(const Dims& D, Lds& L, const uint64_t* __restrict__ word,
                                             const float* __restrict__ pd, const float* __restrict__ ps, int nt) {
    const int ldx = D.nin + 1;
    for (int e = threadIdx.x; e < kTile * D.nin; e += kT) {
        const int s = e / D.nin, k = e % D.nin;
        L.X[s * ldx + k] = s < nt ? feature(D, word[s], pd[s], ps[s], k) : 0.0f;
    }
    __syncthreads();
    const int j = threadIdx.x % kH;
    for (int s = threadIdx.x / kH; s < kTile; s += kT / kH) {
        float z = L.b1[j];
        for (int k = 0; k < D.nin; ++k) z += L.X[s * ldx + k] * L.W1[k * kH + j];
        L.H1[s * kLd + j] = tanhf(z);
    }
    __syncthreads();
    for (int s = threadIdx.x / kH; s < kTile; s += kT / kH) {
        float z = L.b2[j];
#pragma unroll 8
        for (int i = 0; i < kH; ++i) z += L.H1[s * kLd + i] * L.W2[i * kLd + j];
        L.H2[s * kLd + j] = tanhf(z);
    }
    __syncthreads();
    if (threadIdx.x < 2 * kTile) {
        const int s = threadIdx.x / 2, c = threadIdx.x % 2;
        float z = L.b3[c];
#pragma unroll 8
        for (int i = 0; i < kH; ++i) z += L.H2[s * kLd + i] * L.W3[2 * i + c];
        L.O[2 * s + c] = z;
    }
    __syncthreads();
}

// Start vector: psi_i = exp(o0) * sign(o1) for every basis word
__global__ void __launch_bounds__(kT) k_infer(Dims D, const float* __restrict__ P, int64_t n,
                                              const uint64_t* __restrict__ words, float pd, float ps,
                                              double2* __restrict__ out) {
    __shared__ Lds L;
    __shared__ float spd[kTile], sps[kTile];
    stage_params(D, P, L);
    if (threadIdx.x < kTile) {
        spd[threadIdx.x] = pd;
        sps[threadIdx.x] = ps;
    }
    for (int64_t t0 = (int64_t)blockIdx.x * kTile; t0 < n; t0 += (int64_t)gridDim.x * kTile) {
        const int nt = (int)(n - t0 < kTile ? n - t0 : kTile);
        __syncthreads();  // parameters staged / previous tile read
        tile_forward(D, L, words + t0, spd, sps, nt);
        if (threadIdx.x < nt) {
            const double a = exp((double)L.O[2 * threadIdx.x]);
            out[t0 + threadIdx.x] = make_double2(L.O[2 * threadIdx.x + 1] > 0.0f ? a : -a, 0.0);
        }
    }
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Draw the minibatch: set uniform, state ~ |psi|^2 by binary search in the set's CDF.
__global__ void __launch_bounds__(kT) k_sample(int B, uint64_t key, int nsets, const int64_t* __restrict__ set0,
                                               const double* __restrict__ cdf, const uint64_t* __restrict__ words,
                                               const float* __restrict__ ya, const float* __restrict__ ys,
                                               const float* __restrict__ setd, const float* __restrict__ sets,
                                               Samples S) {
    const int s = blockIdx.x * kT + threadIdx.x;
    if (s >= B) return;
    const uint64_t r1 = mix64(key ^ (uint64_t)s * 0xD1B54A32D192ED03ull);
    const uint64_t r2 = mix64(r1);
    const int set = (int)(r1 % (uint64_t)nsets);
    const int64_t a = set0[set], b = set0[set + 1];
    const double u = (double)(r2 >> 11) * (1.0 / 9007199254740992.0);  // [0, 1)
    int64_t lo = a, hi = b - 1;  // first i with cdf[i] > u
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (cdf[mid] > u) hi = mid;
        else lo = mid + 1;
    }
    S.word[s] = words[lo];
    S.pd[s] = setd[set];
    S.ps[s] = sets[set];
    S.ya[s] = ya[lo];
    S.ys[s] = ys[lo];
}

// Forward, loss, back-propagation and the tile's parameter-gradient sums
// (part[tile][param], summed over tiles by k_grad_sum in tile order).
__global__ void __launch_bounds__(kT) k_fwd_bwd(Dims D, const float* __restrict__ P, int B, Samples S,
                                                float* __restrict__ part) {
    __shared__ Lds L;
    __shared__ float G[kTile * 2];  // dL/do per state
    stage_params(D, P, L);
    const int t0 = blockIdx.x * kTile;
    const int nt = B - t0 < kTile ? B - t0 : kTile;
    __syncthreads();
    tile_forward(D, L, S.word + t0, S.pd + t0, S.ps + t0, nt);
    // loss and output deltas (states beyond nt get zero deltas)
    if (threadIdx.x < kTile) {
        const int s = threadIdx.x;
        float g0 = 0.0f, g1 = 0.0f;
        if (s < nt) {
            const float o0 = L.O[2 * s], o1 = L.O[2 * s + 1];
            const float ea = o0 - S.ya[t0 + s];
            const float y = S.ys[t0 + s];
            const float sp = o1 > 0.0f ? o1 + log1pf(expf(-o1)) : log1pf(expf(o1));  // softplus
            S.loss[t0 + s] = ea * ea + sp - y * o1;
            g0 = 2.0f * ea;
            g1 = 1.0f / (1.0f + expf(-o1)) - y;
        }
        G[2 * s] = g0;
        G[2 * s + 1] = g1;
    }
    __syncthreads();
    const int j = threadIdx.x % kH;
    // delta of layer 2: D2 = (1 - H2^2) (W3 G)
    for (int s = threadIdx.x / kH; s < kTile; s += kT / kH) {
        const float h = L.H2[s * kLd + j];
        L.D2[s * kLd + j] = (1.0f - h * h) * (L.W3[2 * j] * G[2 * s] + L.W3[2 * j + 1] * G[2 * s + 1]);
    }
    __syncthreads();
    float* out = part + (int64_t)blockIdx.x * D.n;
    // W3, b3 (H2 is read here for the last time)
    if (threadIdx.x < 2 * kH + 2) {
        const int p = threadIdx.x;
        float acc = 0.0f;
        if (p < 2 * kH) {
            const int i = p / 2, c = p % 2;
            for (int s = 0; s < kTile; ++s) acc += L.H2[s * kLd + i] * G[2 * s + c];
        } else {
            for (int s = 0; s < kTile; ++s) acc += G[2 * s + p - 2 * kH];
        }
        out[D.oW3 + p] = acc;  // oW3 .. ob3+1 are contiguous
    }
    __syncthreads();
    // delta of layer 1 into the H2 slot: D1 = (1 - H1^2) (W2 D2)
    for (int s = threadIdx.x / kH; s < kTile; s += kT / kH) {
        float acc = 0.0f;
#pragma unroll 8
        for (int i = 0; i < kH; ++i) acc += L.W2[j * kLd + i] * L.D2[s * kLd + i];
        const float h = L.H1[s * kLd + j];
        L.H2[s * kLd + j] = (1.0f - h * h) * acc;
    }
    __syncthreads();
    // W2, b2: sum_s H1[s][i] D2[s][j]
    for (int p = threadIdx.x; p < kH * kH + kH; p += kT) {
        float acc = 0.0f;
        if (p < kH * kH) {
            const int i = p / kH, jj = p % kH;
            for (int s = 0; s < kTile; ++s) acc += L.H1[s * kLd + i] * L.D2[s * kLd + jj];
        } else {
            for (int s = 0; s < kTile; ++s) acc += L.D2[s * kLd + p - kH * kH];
        }
        out[D.oW2 + p] = acc;  // oW2 .. ob2+63 are contiguous
    }
    // W1, b1: sum_s X[s][k] D1[s][j]
    const int ldx = D.nin + 1;
    for (int p = threadIdx.x; p < D.nin * kH + kH; p += kT) {
        float acc = 0.0f;
        if (p < D.nin * kH) {
            const int k = p / kH, jj = p % kH;
            for (int s = 0; s < kTile; ++s) acc += L.X[s * ldx + k] * L.H2[s * kLd + jj];
        } else {
            for (int s = 0; s < kTile; ++s) acc += L.H2[s * kLd + p - D.nin * kH];
        }
        out[D.oW1 + p] = acc;  // oW1 .. ob1+63 are contiguous
    }
}

// grad = (sum over tiles of part) / B, tiles in order
__global__ void __launch_bounds__(kT) k_grad_sum(int n, int ntiles, int B, const float* __restrict__ part,
                                                 float* __restrict__ grad) {
    const int p = blockIdx.x * kT + threadIdx.x;
    if (p >= n) return;
    float acc = 0.0f;
    for (int t = 0; t < ntiles; ++t) acc += part[(int64_t)t * n + p];
    grad[p] = acc / (float)B;
}

__global__ void __launch_bounds__(kT) k_adam(int n, float* __restrict__ P, const float* __restrict__ g,
                                             float* __restrict__ m, float* __restrict__ v, float lr, float b1t,
                                             float b2t) {
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i >= n) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    P[i] -= lr * (mi / (1.0f - b1t)) / (sqrtf(vi / (1.0f - b2t)) + eps);
}

// sum of the minibatch losses (one block, fixed order)
__global__ void __launch_bounds__(kT) k_loss_sum(int B, const float* __restrict__ loss, double* __restrict__ out) {
    __shared__ double sh[kT];
    double a = 0.0;
    for (int s = threadIdx.x; s < B; s += kT) a += loss[s];
    sh[threadIdx.x] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < kT; ++i) t += sh[i];
        *out = t / B;
    }
}

struct Buffers {
    DevBuf<uint64_t> word;
    DevBuf<float> pd, ps, ya, ys, loss, part;
    int cap = 0;
    void ensure(int B, int nparams) {
        if (B <= cap) return;
        word.alloc(B);
        pd.alloc(B);
        ps.alloc(B);
        ya.alloc(B);
        ys.alloc(B);
        loss.alloc(B);
        part.alloc((size_t)ceil_div(B, kTile) * nparams);
        cap = B;
    }
    Samples view() { return Samples{word.get(), pd.get(), ps.get(), ya.get(), ys.get(), loss.get()}; }
};

}  // namespace

struct blqh_mlp {
    Dims D{};
    DevBuf<float> P, g, m, v;
    int64_t adam_t = 0;
    Buffers buf;
    DevBuf<double> lsum;
};

blqh_mlp* blqh_mlp_create_impl(int nbits, uint32_t seed) {
    BLQH_REQUIRE(nbits >= 1 && nbits <= 64, BLQH_EINVAL, "nbits must be 1..64");
    auto* M = new blqh_mlp;
    try {
        M->D = make_dims(nbits);
        const int n = M->D.n;
        // Glorot-uniform weights from std::mt19937(seed), zero biases
        std::vector<float> h(n, 0.0f);
        std::mt19937 rng(seed);
        auto fill = [&](int off, int fan_in, int fan_out) {
            const float a = std::sqrt(6.0f / (float)(fan_in + fan_out));
            for (int i = 0; i < fan_in * fan_out; ++i) h[off + i] = (float)((rng() * (1.0 / 4294967296.0)) * 2.0 - 1.0) * a;
        };
        fill(M->D.oW1, M->D.nin, kH);
        fill(M->D.oW2, kH, kH);
        fill(M->D.oW3, kH, 2);
        M->P.alloc(n);
        M->g.alloc(n);
        M->m.alloc(n);
        M->v.alloc(n);
        M->lsum.alloc(1);
        copy_sync(M->P.get(), h.data(), n * 4, hipMemcpyHostToDevice);
        BLQH_HIP(hipMemsetAsync(M->m.get(), 0, n * 4, stream()));
        BLQH_HIP(hipMemsetAsync(M->v.get(), 0, n * 4, stream()));
        BLQH_HIP(hipStreamSynchronize(stream()));
    } catch (...) {
        delete M;
        throw;
    }
    return M;
}

void blqh_mlp_destroy_impl(blqh_mlp* M) { delete M; }
int blqh_mlp_nparams_impl(const blqh_mlp* M) { return M->D.n; }
int blqh_mlp_nbits_impl(const blqh_mlp* M) { return M->D.nbits; }

void blqh_mlp_get_params_impl(const blqh_mlp* M, float* out) {
    copy_sync(out, M->P.get(), M->D.n * 4, hipMemcpyDeviceToHost);
}

void blqh_mlp_set_params_impl(blqh_mlp* M, const float* in) {
    copy_sync(M->P.get(), in, M->D.n * 4, hipMemcpyHostToDevice);
    BLQH_HIP(hipMemsetAsync(M->m.get(), 0, M->D.n * 4, stream()));
    BLQH_HIP(hipMemsetAsync(M->v.get(), 0, M->D.n * 4, stream()));
    BLQH_HIP(hipStreamSynchronize(stream()));
    M->adam_t = 0;
}

// Mean loss and gradient over explicit samples (host arrays).
void blqh_mlp_loss_grad_impl(blqh_mlp* M, int64_t ns, const uint64_t* words, const float* pd, const float* ps,
                             const float* ya, const float* ys, double* loss, float* grad) {
    BLQH_REQUIRE(ns >= 1 && ns <= (1 << 24), BLQH_EINVAL, "sample count out of range");
    hipStream_t st = stream();
    const int B = (int)ns;
    M->buf.ensure(B, M->D.n);
    Samples S = M->buf.view();
    BLQH_HIP(hipMemcpyAsync(S.word, words, B * 8, hipMemcpyHostToDevice, st));
    BLQH_HIP(hipMemcpyAsync(S.pd, pd, B * 4, hipMemcpyHostToDevice, st));
    BLQH_HIP(hipMemcpyAsync(S.ps, ps, B * 4, hipMemcpyHostToDevice, st));
    BLQH_HIP(hipMemcpyAsync(S.ya, ya, B * 4, hipMemcpyHostToDevice, st));
    BLQH_HIP(hipMemcpyAsync(S.ys, ys, B * 4, hipMemcpyHostToDevice, st));
    const int nt = (int)ceil_div(B, kTile);
    hipLaunchKernelGGL(k_fwd_bwd, dim3(nt), dim3(kT), 0, st, M->D, M->P.get(), B, S, M->buf.part.get());
    hipLaunchKernelGGL(k_grad_sum, dim3(ceil_div(M->D.n, kT)), dim3(kT), 0, st, M->D.n, nt, B, M->buf.part.get(),
                       M->g.get());
    hipLaunchKernelGGL(k_loss_sum, dim3(1), dim3(kT), 0, st, B, S.loss, M->lsum.get());
    BLQH_HIP(hipGetLastError());
    if (grad) BLQH_HIP(hipMemcpyAsync(grad, M->g.get(), M->D.n * 4, hipMemcpyDeviceToHost, st));
    double l = 0;
    BLQH_HIP(hipMemcpyAsync(&l, M->lsum.get(), 8, hipMemcpyDeviceToHost, st));
    BLQH_HIP(hipStreamSynchronize(st));
    if (loss) *loss = l;
}

// Train on ground states: set k = (basis, d, Delta_SAS, psi).
void blqh_mlp_train_impl(blqh_mlp* M, int nsets, const blqh_basis* const* bases, const double* params,
                         const double* const* psi, const blqh_mlp_train_opts& o, blqh_mlp_train_result& r) {
    auto t0 = std::chrono::steady_clock::now();
    BLQH_REQUIRE(nsets >= 1 && bases && params && psi, BLQH_EINVAL, "bad training sets");
    BLQH_REQUIRE(o.steps >= 1 && o.batch >= 1 && o.batch <= (1 << 22) && o.lr > 0, BLQH_EINVAL,
                 "bad training options");
    hipStream_t st = stream();
    // labels: gauge-fixed real ground states, |psi|^2 CDFs
    std::vector<int64_t> set0(nsets + 1, 0);
    for (int k = 0; k < nsets; ++k) {
        BLQH_REQUIRE(bases[k] && psi[k], BLQH_EINVAL, "null training set");
        BLQH_REQUIRE(bases[k]->geo.S * bases[k]->geo.nphi <= M->D.nbits, BLQH_EINVAL, "basis words wider than the network input");
        set0[k + 1] = set0[k] + bases[k]->n;
    }
    const int64_t N = set0[nsets];
    std::vector<uint64_t> words(N);
    std::vector<float> ya(N), ys(N);
    std::vector<double> cdf(N);
    std::vector<float> sd(nsets), ss(nsets);
    for (int k = 0; k < nsets; ++k) {
        const int64_t n = bases[k]->n, a = set0[k];
        std::memcpy(words.data() + a, bases[k]->h_id.data(), n * 8);
        const double2* z = (const double2*)psi[k];
        int64_t imax = 0;
        double amax = -1;
        for (int64_t i = 0; i < n; ++i) {
            const double m2 = z[i].x * z[i].x + z[i].y * z[i].y;
            if (m2 > amax) { amax = m2; imax = i; }
        }
        BLQH_REQUIRE(amax > 0, BLQH_EINVAL, "zero training state");
        const double ph = std::atan2(z[imax].y, z[imax].x);
        const double c = std::cos(ph), s = std::sin(ph);
        double tot = 0;
        for (int64_t i = 0; i < n; ++i) {
            const double re = z[i].x * c + z[i].y * s;  // Re(psi e^{-i ph})
            const double m2 = z[i].x * z[i].x + z[i].y * z[i].y;
            ya[a + i] = (float)std::log(std::max(std::sqrt(m2), 1e-30));
            ys[a + i] = re > 0 ? 1.0f : 0.0f;
            tot += m2;
            cdf[a + i] = tot;
        }
        for (int64_t i = 0; i < n; ++i) cdf[a + i] /= tot;
        cdf[a + n - 1] = 1.0;
        sd[k] = (float)params[2 * k];
        ss[k] = (float)params[2 * k + 1];
    }
    DevBuf<int64_t> dset0(nsets + 1);
    DevBuf<uint64_t> dwords(N);
    DevBuf<float> dya(N), dys(N), dsd(nsets), dss(nsets);
    DevBuf<double> dcdf(N);
    BLQH_HIP(hipMemcpyAsync(dset0.get(), set0.data(), (nsets + 1) * 8, hipMemcpyHostToDevice, st));
    BLQH_HIP(hipMemcpyAsync(dwords.get(), words.data(), N * 8, hipMemcpyHostToDevice, st));
    BLQH_HIP(hipMemcpyAsync(dya.get(), ya.data(), N * 4, hipMemcpyHostToDevice, st));
    BLQH_HIP(hipMemcpyAsync(dys.get(), ys.data(), N * 4, hipMemcpyHostToDevice, st));
    BLQH_HIP(hipMemcpyAsync(dcdf.get(), cdf.data(), N * 8, hipMemcpyHostToDevice, st));
    BLQH_HIP(hipMemcpyAsync(dsd.get(), sd.data(), nsets * 4, hipMemcpyHostToDevice, st));
    BLQH_HIP(hipMemcpyAsync(dss.get(), ss.data(), nsets * 4, hipMemcpyHostToDevice, st));
    const int B = o.batch;
    M->buf.ensure(B, M->D.n);
    Samples S = M->buf.view();
    const int gb = (int)ceil_div(B, kT), gp = (int)ceil_div(M->D.n, kT), ntile = (int)ceil_div(B, kTile);
    const double b1 = 0.9, b2 = 0.999;
    std::vector<double> losses;
    auto record = [&]() {
        hipLaunchKernelGGL(k_loss_sum, dim3(1), dim3(kT), 0, st, B, S.loss, M->lsum.get());
        double l = 0;
        BLQH_HIP(hipMemcpyAsync(&l, M->lsum.get(), 8, hipMemcpyDeviceToHost, st));
        BLQH_HIP(hipStreamSynchronize(st));
        losses.push_back(l);
    };
    const int tail = std::max(1, std::min(o.steps, 20));  // last steps averaged for the reported loss
    double last = 0;
    for (int step = 0; step < o.steps; ++step) {
        const uint64_t key = ((uint64_t)o.seed << 32) ^ (uint64_t)(M->adam_t + 1) * 0x9E3779B97F4A7C15ull;
        hipLaunchKernelGGL(k_sample, dim3(gb), dim3(kT), 0, st, B, key, nsets, dset0.get(), dcdf.get(),
                           dwords.get(), dya.get(), dys.get(), dsd.get(), dss.get(), S);
        hipLaunchKernelGGL(k_fwd_bwd, dim3(ntile), dim3(kT), 0, st, M->D, M->P.get(), B, S, M->buf.part.get());
        if (step == 0) record();
        if (step >= o.steps - tail) {
            record();
            last += losses.back();
        }
        hipLaunchKernelGGL(k_grad_sum, dim3(gp), dim3(kT), 0, st, M->D.n, ntile, B, M->buf.part.get(), M->g.get());
        ++M->adam_t;
        const float lr = o.lr * (o.lr_final > 0 ? (float)std::pow(o.lr_final / o.lr, (double)step / std::max(1, o.steps - 1)) : 1.0f);
        hipLaunchKernelGGL(k_adam, dim3(gp), dim3(kT), 0, st, M->D.n, M->P.get(), M->g.get(), M->m.get(),
                           M->v.get(), lr, (float)std::pow(b1, (double)M->adam_t), (float)std::pow(b2, (double)M->adam_t));
        BLQH_HIP(hipGetLastError());
    }
    BLQH_HIP(hipStreamSynchronize(st));
    r.loss_first = losses.front();
    r.loss_last = last / tail;
    r.steps = o.steps;
    r.samples = N;
    r.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Start vector (normalised, real) for a basis at (d, Delta_SAS).
void blqh_mlp_start_vector_impl(const blqh_mlp* M, const blqh_basis* B, double d, double dsas, double* out) {
    BLQH_REQUIRE(B->geo.S * B->geo.nphi <= M->D.nbits, BLQH_EINVAL, "basis words wider than the network input");
    hipStream_t st = stream();
    const int64_t n = B->n;
    DevBuf<double2> v(std::max<int64_t>(n, 1));
    const int g = (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, kTile), 4096));
    hipLaunchKernelGGL(k_infer, dim3(g), dim3(kT), 0, st, M->D, M->P.get(), n, B->id.get(), (float)d, (float)dsas,
                       v.get());
    BLQH_HIP(hipGetLastError());
    BLQH_HIP(hipMemcpyAsync(out, v.get(), n * sizeof(double2), hipMemcpyDeviceToHost, st));
    BLQH_HIP(hipStreamSynchronize(st));
    double s = 0;
    for (int64_t i = 0; i < n; ++i) s += out[2 * i] * out[2 * i];
    BLQH_REQUIRE(s > 0 && std::isfinite(s), BLQH_EINTERNAL, "network start vector is zero or not finite");
    s = 1.0 / std::sqrt(s);
    for (int64_t i = 0; i < 2 * n; ++i) out[i] *= s;
}
