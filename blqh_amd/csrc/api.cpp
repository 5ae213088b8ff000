// api.cpp -- the C ABI (include/blqh.h): argument checks, error codes and
// the thread-local error message; no exception crosses the boundary.
#include <hip/hip_runtime.h>

#include <cstring>
#include <list>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "engine.hpp"

void blqh_basis_build_impl(const blqh_sector& s, blqh_basis* B);
void blqh_hamil_build_impl(const blqh_basis* B, const blqh_params& p, blqh_hamil* H, int terms);
void blqh_hamil_combine_impl(const blqh_hamil* base, const blqh_hamil* tun, double dsas, blqh_hamil* H);
void blqh_spmv_device_impl(blqh_hamil* H, const void* x, void* y);
void blqh_lanczos_impl(blqh_hamil* H, const blqh_lanczos_opts& o, blqh_lanczos_result& r, double* psi, double* alpha,
                       double* beta);
void blqh_lanczos_noalloc_impl(blqh_hamil* H, const blqh_lanczos_opts& o, blqh_lanczos_result& r, double* psi,
                               double* alpha, double* beta);
blqh_batch* blqh_batch_create_impl(blqh_hamil* const* hs, int count, int lanes);
void blqh_batch_destroy_impl(blqh_batch* b);
void blqh_batch_rebind_impl(blqh_batch* b, blqh_hamil* const* hs, int count);
blqh_batch* blqh_batch_create_scaled_impl(const blqh_hamil* const* base, const blqh_hamil* const* tun, int nsec,
                                          const double* scales, int nscale);
void blqh_batch_rebind_scaled_impl(blqh_batch* b, const blqh_hamil* const* base, const blqh_hamil* const* tun,
                                   int nsec);
int blqh_batch_lanes_impl(const blqh_batch* b);
void blqh_batch_set_lanes_impl(blqh_batch* b, int lanes);
void blqh_lanczos_reconstruct_impl(blqh_hamil* H, int m, uint32_t seed, const double* start, const double* alpha,
                                   const double* beta, int ns, const double* coefs, double* psi, double* erq,
                                   double* dev);
void blqh_tridiag_vectors_impl(int m, const double* alpha, const double* offdiag, int k0, int count, double* vecs);
void blqh_batch_lanczos_impl(blqh_batch* b, const blqh_lanczos_opts& o, const uint32_t* seeds,
                             const double* const* starts, blqh_lanczos_result* res, double* const* psi_out,
                             bool cont, double excited);
blqh_mlp* blqh_mlp_create_impl(int nbits, uint32_t seed);
void blqh_mlp_destroy_impl(blqh_mlp* M);
int blqh_mlp_nparams_impl(const blqh_mlp* M);
void blqh_mlp_get_params_impl(const blqh_mlp* M, float* out);
void blqh_mlp_set_params_impl(blqh_mlp* M, const float* in);
void blqh_mlp_loss_grad_impl(blqh_mlp* M, int64_t ns, const uint64_t* words, const float* pd, const float* ps,
                             const float* ya, const float* ys, double* loss, float* grad);
void blqh_mlp_train_impl(blqh_mlp* M, int nsets, const blqh_basis* const* bases, const double* params,
                         const double* const* psi, const blqh_mlp_train_opts& o, blqh_mlp_train_result& r);
void blqh_mlp_start_vector_impl(const blqh_mlp* M, const blqh_basis* B, double d, double dsas, double* out);
void blqh_tridiag_spectrum_impl(int m, const double* alpha, const double* offdiag, double* eig, double* first,
                                double* vec0);
int blqh_spectrum_distinct_impl(int m, const double* eig, int kmax, double gap, double* out);
void blqh_record_save_impl(const char* path, int64_t nhilbert, int m, uint32_t seed, const double* eig,
                           const double* beta, const double* alpha, const double* psi0, const double* psin0);
void blqh_record_load_impl(const char* path, int cap, int64_t* nhilbert, int* m, uint32_t* seed, double* eig,
                           double* beta, double* alpha, double* psi0, double* psin0);
void blqh_observables_impl(const blqh_basis* B, const double* psi_host, blqh_observables& out);
void blqh_dense_ground_state_impl(const blqh_hamil* H, double* e0, double* e_printed, double* vec);
void blqh_device_copy_impl(const void* src, void* dst, int64_t bytes);
namespace blqh {
void noalloc_setup(int device);
void start_vector(int64_t n, uint32_t seed, std::vector<double2>& v);
double tridiag_lowest(int m, const double* a, const double* b, double* vec);
}  // namespace blqh

namespace {
thread_local std::string g_err;
std::mutex g_init_mu;
blqh::Runtime g_rt;
bool g_rt_ok = false;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        return BLQH_OK;
    } catch (const blqh::Error& e) {
        return fail(e.code, e.what());
    } catch (const std::bad_alloc&) {
        return fail(BLQH_ENOMEM, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(BLQH_EINTERNAL, e.what());
    }
}

void init_device(int dev) {
    std::lock_guard<std::mutex> lk(g_init_mu);
    if (g_rt_ok) {
        BLQH_REQUIRE(dev < 0 || dev == g_rt.device, BLQH_EINVAL, "library already bound to another device");
        return;
    }
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) throw blqh::Error(BLQH_EDEVICE, "no HIP device available");
    if (dev < 0) dev = 0;
    BLQH_REQUIRE(dev < count, BLQH_EINVAL, "device index out of range");
    BLQH_HIP(hipSetDevice(dev));
    hipDeviceProp_t prop;
    BLQH_HIP(hipGetDeviceProperties(&prop, dev));
    g_rt.device = dev;
    g_rt.num_cus = prop.multiProcessorCount;
    blqh::noalloc_setup(dev);  // device address of the allocation-free solver's static scratch
    g_rt_ok = true;
}

// Library streams: one per host thread (non-blocking, default priority),
// taken on the thread's first library call that needs one and handed back to
// a free list when the thread ends.  Streams are never destroyed: the HIP
// runtime may already be gone when the main thread's thread-locals end.
std::mutex g_streams_mu;
std::vector<hipStream_t>* g_free_streams = new std::vector<hipStream_t>;  // never destroyed
struct ThreadStream {
    hipStream_t s = nullptr;
    ~ThreadStream() {
        if (!s) return;
        std::lock_guard<std::mutex> lk(g_streams_mu);
        g_free_streams->push_back(s);
        s = nullptr;
    }
};
thread_local ThreadStream t_stream;
}  // namespace

namespace blqh {
Runtime& runtime() {
    if (!g_rt_ok) init_device(-1);
    return g_rt;
}
hipStream_t stream() {
    if (t_stream.s) return t_stream.s;
    const Runtime& rt = runtime();
    BLQH_HIP(hipSetDevice(rt.device));  // a new host thread starts on device 0
    std::lock_guard<std::mutex> lk(g_streams_mu);
    if (!g_free_streams->empty()) {
        t_stream.s = g_free_streams->back();
        g_free_streams->pop_back();
    } else {
        BLQH_HIP(hipStreamCreateWithFlags(&t_stream.s, hipStreamNonBlocking));
    }
    return t_stream.s;
}

namespace {
// Released device blocks, oldest first, with an event recorded on the
// releasing thread's stream: a block goes to a new owner once every kernel
// queued before its release has finished (hipEventSynchronize on that event),
// so another thread's queued work is not waited for.  Every size is kept
// (hipFree waits for the whole device, which would stall a solve on another
// thread); at most a quarter of device memory and kMaxBlocks blocks.
struct CachedBlock {
    size_t bytes;
    void* p;
    hipEvent_t ev;
};
struct BlockCache {
    std::mutex mu;
    std::list<CachedBlock> blocks;              // oldest first
    std::multimap<size_t, std::list<CachedBlock>::iterator> by_size;
    std::vector<hipEvent_t> spare_events;
    size_t bytes = 0, cap = 0;
};
constexpr size_t kMaxBlocks = 4096;
BlockCache& block_cache() {
    static BlockCache* c = new BlockCache;  // never destroyed: no hipFree after the runtime shuts down
    return *c;
}
void cache_erase(BlockCache& c, std::list<CachedBlock>::iterator it) {  // caller holds c.mu
    auto r = c.by_size.equal_range(it->bytes);
    for (auto m = r.first; m != r.second; ++m)
        if (m->second == it) {
            c.by_size.erase(m);
            break;
        }
    c.bytes -= it->bytes;
    c.spare_events.push_back(it->ev);
    c.blocks.erase(it);
}
void cache_drop_oldest(BlockCache& c) {  // caller holds c.mu
    void* p = c.blocks.front().p;
    cache_erase(c, c.blocks.begin());
    (void)hipFree(p);
}
void cache_drop_all(BlockCache& c) {  // caller holds c.mu
    while (!c.blocks.empty()) cache_drop_oldest(c);
}
}  // namespace

void* dev_alloc(size_t bytes) {
    (void)stream();  // binds the calling thread to the library's device
    BlockCache& c = block_cache();
    void* hit = nullptr;
    hipEvent_t ev = nullptr;
    {
        std::lock_guard<std::mutex> lk(c.mu);
        auto m = c.by_size.find(bytes);
        if (m != c.by_size.end()) {
            hit = m->second->p;
            ev = m->second->ev;
            cache_erase(c, m->second);  // the event goes to the spares only after the wait below
            c.spare_events.pop_back();
        }
    }
    if (hit) {
        const hipError_t e = hipEventSynchronize(ev);
        {
            std::lock_guard<std::mutex> lk(c.mu);
            c.spare_events.push_back(ev);
        }
        BLQH_HIP(e);
        return hit;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, std::max<size_t>(bytes, 1));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        {
            std::lock_guard<std::mutex> lk(c.mu);
            cache_drop_all(c);
        }
        e = hipMalloc(&p, std::max<size_t>(bytes, 1));
    }
    if (e != hipSuccess)
        throw Error(BLQH_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
    return p;
}

void dev_free(void* p, size_t bytes) noexcept {
    if (!p) return;
    BlockCache& c = block_cache();
    if (g_rt_ok) {
        hipStream_t st = nullptr;
        try {
            st = stream();
        } catch (...) {
            st = nullptr;
        }
        std::lock_guard<std::mutex> lk(c.mu);
        if (c.cap == 0) {
            size_t fr = 0, tot = 0;
            if (hipMemGetInfo(&fr, &tot) == hipSuccess) c.cap = tot / 4;
        }
        hipEvent_t ev = nullptr;
        if (!c.spare_events.empty()) {
            ev = c.spare_events.back();
            c.spare_events.pop_back();
        } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
            ev = nullptr;
        }
        if (st && ev && bytes <= c.cap && hipEventRecord(ev, st) == hipSuccess) {
            while ((c.bytes + bytes > c.cap || c.blocks.size() >= kMaxBlocks) && !c.blocks.empty())
                cache_drop_oldest(c);
            c.blocks.push_back({bytes, p, ev});
            c.by_size.emplace(bytes, std::prev(c.blocks.end()));
            c.bytes += bytes;
            return;
        }
        if (ev) c.spare_events.push_back(ev);
    }
    (void)hipFree(p);
}

size_t dev_cached_bytes() {
    BlockCache& c = block_cache();
    std::lock_guard<std::mutex> lk(c.mu);
    return c.bytes;
}

size_t dev_release_cached() {
    BlockCache& c = block_cache();
    std::lock_guard<std::mutex> lk(c.mu);
    const size_t b = c.bytes;
    cache_drop_all(c);
    return b;
}
}  // namespace blqh

extern "C" {

const char* blqh_last_error(void) { return g_err.c_str(); }

int blqh_init(int device) {
    return guarded([&] { init_device(device); });
}

int blqh_basis_create(const blqh_sector* s, blqh_basis** out) {
    if (!s || !out) return fail(BLQH_EINVAL, "null argument");
    *out = nullptr;
    return guarded([&] {
        blqh::runtime();
        auto* B = new blqh_basis;
        try {
            blqh_basis_build_impl(*s, B);
        } catch (...) {
            delete B;
            throw;
        }
        *out = B;
    });
}

int64_t blqh_basis_size(const blqh_basis* b) {
    if (!b) return fail(BLQH_EINVAL, "null basis");
    return b->n;
}

int blqh_basis_export(const blqh_basis* b, uint64_t* id, int16_t* period) {
    if (!b) return fail(BLQH_EINVAL, "null basis");
    if (id) std::memcpy(id, b->h_id.data(), b->h_id.size() * sizeof(uint64_t));
    if (period) std::memcpy(period, b->h_period.data(), b->h_period.size() * sizeof(int16_t));
    return BLQH_OK;
}

void blqh_basis_destroy(blqh_basis* b) { delete b; }

int blqh_hamil_build(const blqh_basis* b, const blqh_params* p, blqh_hamil** out) {
    if (!b || !p || !out) return fail(BLQH_EINVAL, "null argument");
    *out = nullptr;
    return guarded([&] {
        auto* H = new blqh_hamil;
        try {
            blqh_hamil_build_impl(b, *p, H, 3);
        } catch (...) {
            delete H;
            throw;
        }
        *out = H;
    });
}

int blqh_hamil_build_terms(const blqh_basis* b, const blqh_params* p, int32_t terms, blqh_hamil** out) {
    if (!b || !p || !out) return fail(BLQH_EINVAL, "null argument");
    *out = nullptr;
    return guarded([&] {
        auto* H = new blqh_hamil;
        try {
            blqh_hamil_build_impl(b, *p, H, terms);
        } catch (...) {
            delete H;
            throw;
        }
        *out = H;
    });
}

int blqh_hamil_prepare(blqh_hamil* h) {
    if (!h) return fail(BLQH_EINVAL, "null hamiltonian");
    return guarded([&] { blqh::ensure_panels(h); });
}

int blqh_hamil_combine(const blqh_hamil* base, const blqh_hamil* tun, double delta_sas, blqh_hamil** out) {
    if (!base || !tun || !out) return fail(BLQH_EINVAL, "null argument");
    *out = nullptr;
    return guarded([&] {
        auto* H = new blqh_hamil;
        try {
            blqh_hamil_combine_impl(base, tun, delta_sas, H);
        } catch (...) {
            delete H;
            throw;
        }
        *out = H;
    });
}

int blqh_hamil_info(const blqh_hamil* h, int64_t* nrows, int64_t* nnz) {
    if (!h) return fail(BLQH_EINVAL, "null hamiltonian");
    if (nrows) *nrows = h->n;
    if (nnz) *nnz = h->nnz;
    return BLQH_OK;
}

int blqh_hamil_export(const blqh_hamil* h, int64_t* rowptr, int32_t* col, double* val) {
    if (!h) return fail(BLQH_EINVAL, "null hamiltonian");
    return guarded([&] {
        hipStream_t st = blqh::stream();
        if (rowptr) BLQH_HIP(hipMemcpyAsync(rowptr, h->rowptr.get(), (h->n + 1) * 8, hipMemcpyDeviceToHost, st));
        if (col && h->nnz) BLQH_HIP(hipMemcpyAsync(col, h->col.get(), h->nnz * 4, hipMemcpyDeviceToHost, st));
        if (val && h->nnz) BLQH_HIP(hipMemcpyAsync(val, h->val.get(), h->nnz * 16, hipMemcpyDeviceToHost, st));
        BLQH_HIP(hipStreamSynchronize(st));
    });
}

void blqh_hamil_destroy(blqh_hamil* h) { delete h; }

int blqh_hamil_get_spmv(const blqh_hamil* h, int32_t* kind, int32_t* lanes, int32_t* threads, int32_t* wpc,
                        int32_t* rows) {
    if (!h) return fail(BLQH_EINVAL, "null hamiltonian");
    if (kind) *kind = h->spmv_kind;
    if (lanes) *lanes = h->spmv_lpr;
    if (threads) *threads = h->spmv_threads;
    if (wpc) *wpc = h->spmv_wgs_per_cu;
    if (rows) *rows = h->spmv_rows;
    return BLQH_OK;
}

int blqh_hamil_set_spmv(blqh_hamil* h, int32_t kind, int32_t lanes, int32_t threads, int32_t wpc, int32_t rows) {
    if (!h) return fail(BLQH_EINVAL, "null hamiltonian");
    if (kind >= 1) {  // the panel kinds need the SpMV formats
        const int rc = guarded([&] { blqh::ensure_panels(h); });
        if (rc != BLQH_OK) return rc;
    }
    bool ok = false;
    if (kind == -1) ok = true;
    if (kind == 0) ok = lanes == 4 || lanes == 8 || lanes == 16 || lanes == 32 || lanes == 64;
    if (kind == 1)
        ok = h->pan.built && (lanes == 4 || lanes == 8 || lanes == 16) && (threads == 512 || threads == 1024) &&
             (wpc == 1 || wpc == 2) && (rows == 2 || rows == 4);
    if (kind == 2)
        ok = h->pan.built && lanes == 64 && (threads == 512 || threads == 1024) && wpc >= 1 && wpc <= 4 &&
             (rows == 1 || rows == 2 || rows == 4 || rows == 8);
    if (kind == 3)
        ok = h->pan.built && lanes == 64 && (threads == 256 || threads == 512 || threads == 1024) && wpc >= 1 &&
             wpc <= 4 && (rows == 1 || rows == 2 || rows == 4 || rows == 8);
    if (kind == 5)
        ok = h->pan.built && (lanes == 2 || lanes == 4 || lanes == 8) && (threads == 512 || threads == 1024) &&
             wpc >= 1 && wpc <= 4 && (rows == 1 || rows == 2 || rows == 4) && lanes * rows <= 8;
    if (kind == 4) {
        // the column-tile layout is built here on first use (one workgroup per CU)
        int rc = guarded([&] { ok = blqh::build_tiles(h, blqh::runtime().num_cus); });
        if (rc != BLQH_OK) return rc;
        if (ok) {
            lanes = 64;
            threads = 1024;
            wpc = 1;
            rows = h->til.R;
        }
    }
    if (!ok) return fail(BLQH_EINVAL, "unsupported SpMV variant");
    h->spmv_kind = kind;
    h->spmv_lpr = kind < 0 ? 0 : lanes;
    h->spmv_threads = kind >= 1 ? threads : 256;
    h->spmv_wgs_per_cu = kind >= 1 ? wpc : 0;
    h->spmv_rows = kind >= 1 ? rows : 0;
    return BLQH_OK;
}

int blqh_spmv(const blqh_hamil* h, const double* x, double* y) {
    if (!h || !x || !y) return fail(BLQH_EINVAL, "null argument");
    return guarded([&] {
        hipStream_t st = blqh::stream();
        blqh::DevBuf<double2> dx(h->n), dy(h->n);
        BLQH_HIP(hipMemcpyAsync(dx.get(), x, h->n * 16, hipMemcpyHostToDevice, st));
        blqh_spmv_device_impl(const_cast<blqh_hamil*>(h), dx.get(), dy.get());
        BLQH_HIP(hipMemcpyAsync(y, dy.get(), h->n * 16, hipMemcpyDeviceToHost, st));
        BLQH_HIP(hipStreamSynchronize(st));
    });
}

int blqh_spmv_device(const blqh_hamil* h, const void* x, void* y) {
    if (!h || !x || !y) return fail(BLQH_EINVAL, "null argument");
    return guarded([&] { blqh_spmv_device_impl(const_cast<blqh_hamil*>(h), x, y); });
}

int blqh_lanczos(const blqh_hamil* h, const blqh_lanczos_opts* o, blqh_lanczos_result* r, double* psi, double* alpha,
                 double* beta) {
    if (!h || !o || !r) return fail(BLQH_EINVAL, "null argument");
    return guarded([&] {
        std::memset(r, 0, sizeof(*r));
        blqh_lanczos_impl(const_cast<blqh_hamil*>(h), *o, *r, psi, alpha, beta);
    });
}

int blqh_lanczos_noalloc(const blqh_hamil* h, const blqh_lanczos_opts* o, blqh_lanczos_result* r, double* psi,
                         double* alpha, double* beta) {
    if (!h || !o || !r) return fail(BLQH_EINVAL, "null argument");
    return guarded([&] {
        std::memset(r, 0, sizeof(*r));
        blqh_lanczos_noalloc_impl(const_cast<blqh_hamil*>(h), *o, *r, psi, alpha, beta);
    });
}

int blqh_lanczos_reconstruct(const blqh_hamil* h, int32_t m, uint32_t seed, const double* start,
                             const double* alpha, const double* beta, int32_t ns, const double* coefs, double* psi,
                             double* erq, double* max_dev) {
    if (!h || !coefs) return fail(BLQH_EINVAL, "null argument");
    return guarded([&] {
        blqh_lanczos_reconstruct_impl(const_cast<blqh_hamil*>(h), m, seed, start, alpha, beta, ns, coefs, psi, erq,
                                      max_dev);
    });
}

int blqh_batch_create(blqh_hamil* const* hs, int32_t count, int32_t lanes, blqh_batch** out) {
    if (!hs || !out || count < 1) return fail(BLQH_EINVAL, "bad argument");
    *out = nullptr;
    return guarded([&] { *out = blqh_batch_create_impl(hs, count, lanes); });
}

int blqh_batch_rebind(blqh_batch* b, blqh_hamil* const* hs, int32_t count) {
    if (!b || !hs) return fail(BLQH_EINVAL, "null argument");
    return guarded([&] { blqh_batch_rebind_impl(b, hs, count); });
}

int blqh_batch_create_scaled(const blqh_hamil* const* base, const blqh_hamil* const* tun, int32_t nsector,
                             const double* scales, int32_t nscale, blqh_batch** out) {
    if (!base || !tun || !scales || !out || nsector < 1) return fail(BLQH_EINVAL, "bad argument");
    *out = nullptr;
    return guarded([&] { *out = blqh_batch_create_scaled_impl(base, tun, nsector, scales, nscale); });
}

int blqh_batch_rebind_scaled(blqh_batch* b, const blqh_hamil* const* base, const blqh_hamil* const* tun,
                             int32_t nsector) {
    if (!b || !base || !tun) return fail(BLQH_EINVAL, "null argument");
    return guarded([&] { blqh_batch_rebind_scaled_impl(b, base, tun, nsector); });
}

int blqh_batch_lanczos(blqh_batch* b, const blqh_lanczos_opts* o, const uint32_t* seeds, const double* const* starts,
                       blqh_lanczos_result* res, double* const* psi) {
    if (!b || !o || !res) return fail(BLQH_EINVAL, "null argument");
    return guarded([&] { blqh_batch_lanczos_impl(b, *o, seeds, starts, res, psi, false, 0.0); });
}

int blqh_batch_lanczos_continue(blqh_batch* b, const blqh_lanczos_opts* o, const uint32_t* seeds,
                                double excited_mix, blqh_lanczos_result* res, double* const* psi) {
    if (!b || !o || !res) return fail(BLQH_EINVAL, "null argument");
    return guarded([&] { blqh_batch_lanczos_impl(b, *o, seeds, nullptr, res, psi, true, excited_mix); });
}

int blqh_batch_get_lanes(const blqh_batch* b, int32_t* lanes) {
    if (!b || !lanes) return fail(BLQH_EINVAL, "null argument");
    *lanes = blqh_batch_lanes_impl(b);
    return BLQH_OK;
}

int blqh_batch_set_lanes(blqh_batch* b, int32_t lanes) {
    if (!b) return fail(BLQH_EINVAL, "null batch");
    return guarded([&] { blqh_batch_set_lanes_impl(b, lanes); });
}

void blqh_batch_destroy(blqh_batch* b) {
    if (b) blqh_batch_destroy_impl(b);
}

int blqh_mlp_create(int32_t nbits, uint32_t seed, blqh_mlp** out) {
    if (!out) return fail(BLQH_EINVAL, "null argument");
    *out = nullptr;
    return guarded([&] {
        blqh::runtime();
        *out = blqh_mlp_create_impl(nbits, seed);
    });
}

int blqh_mlp_nparams(const blqh_mlp* m) {
    if (!m) return fail(BLQH_EINVAL, "null network");
    return blqh_mlp_nparams_impl(m);
}

int blqh_mlp_get_params(const blqh_mlp* m, float* out) {
    if (!m || !out) return fail(BLQH_EINVAL, "null argument");
    return guarded([&] { blqh_mlp_get_params_impl(m, out); });
}

int blqh_mlp_set_params(blqh_mlp* m, const float* in) {
    if (!m || !in) return fail(BLQH_EINVAL, "null argument");
    return guarded([&] { blqh_mlp_set_params_impl(m, in); });
}

int blqh_mlp_loss_grad(blqh_mlp* m, int64_t count, const uint64_t* words, const float* d, const float* dsas,
                       const float* ya, const float* ys, double* loss, float* grad) {
    if (!m || !words || !d || !dsas || !ya || !ys) return fail(BLQH_EINVAL, "null argument");
    return guarded([&] { blqh_mlp_loss_grad_impl(m, count, words, d, dsas, ya, ys, loss, grad); });
}

int blqh_mlp_train(blqh_mlp* m, int32_t nsets, const blqh_basis* const* bases, const double* params,
                   const double* const* psi, const blqh_mlp_train_opts* o, blqh_mlp_train_result* r) {
    if (!m || !o || !r) return fail(BLQH_EINVAL, "null argument");
    return guarded([&] {
        std::memset(r, 0, sizeof(*r));
        blqh_mlp_train_impl(m, nsets, bases, params, psi, *o, *r);
    });
}

int blqh_mlp_start_vector(const blqh_mlp* m, const blqh_basis* b, double d, double dsas, double* out) {
    if (!m || !b || !out) return fail(BLQH_EINVAL, "null argument");
    return guarded([&] { blqh_mlp_start_vector_impl(m, b, d, dsas, out); });
}

void blqh_mlp_destroy(blqh_mlp* m) {
    if (m) blqh_mlp_destroy_impl(m);
}

int blqh_compute_observables(const blqh_basis* b, const double* psi, blqh_observables* out) {
    if (!b || !psi || !out) return fail(BLQH_EINVAL, "null argument");
    return guarded([&] { blqh_observables_impl(b, psi, *out); });
}

int blqh_tridiag_spectrum(int32_t m, const double* alpha, const double* offdiag, double* eig, double* first,
                          double* vec0) {
    if (m < 1 || !alpha || (m > 1 && !offdiag)) return fail(BLQH_EINVAL, "bad argument");
    return guarded([&] { blqh_tridiag_spectrum_impl(m, alpha, offdiag, eig, first, vec0); });
}

int blqh_tridiag_vectors(int32_t m, const double* alpha, const double* offdiag, int32_t k0, int32_t count,
                         double* vecs) {
    if (m < 1 || !alpha || !vecs || (m > 1 && !offdiag)) return fail(BLQH_EINVAL, "bad argument");
    return guarded([&] { blqh_tridiag_vectors_impl(m, alpha, offdiag, k0, count, vecs); });
}

int blqh_spectrum_distinct(int32_t m, const double* eig, int32_t kmax, double gap, double* out) {
    if (m < 0 || kmax < 0 || (m > 0 && (!eig || !out))) return fail(BLQH_EINVAL, "bad argument");
    return blqh_spectrum_distinct_impl(m, eig, kmax, gap, out);
}

int blqh_record_save(const char* path, int64_t nhilbert, int32_t m, uint32_t seed, const double* eig,
                     const double* beta, const double* alpha, const double* psi0, const double* psin0) {
    if (!path || m < 1 || !eig || !beta || !alpha || !psi0 || !psin0) return fail(BLQH_EINVAL, "bad argument");
    return guarded([&] { blqh_record_save_impl(path, nhilbert, m, seed, eig, beta, alpha, psi0, psin0); });
}

int blqh_record_load(const char* path, int32_t cap, int64_t* nhilbert, int32_t* m, uint32_t* seed, double* eig,
                     double* beta, double* alpha, double* psi0, double* psin0) {
    if (!path || !nhilbert || !m || !seed || cap < 1) return fail(BLQH_EINVAL, "bad argument");
    return guarded([&] { blqh_record_load_impl(path, cap, nhilbert, m, seed, eig, beta, alpha, psi0, psin0); });
}

int blqh_dense_ground_state(const blqh_hamil* h, double* e0, double* e_printed, double* vec) {
    if (!h) return fail(BLQH_EINVAL, "null hamiltonian");
    return guarded([&] { blqh_dense_ground_state_impl(h, e0, e_printed, vec); });
}

int blqh_device_copy(const void* src, void* dst, int64_t bytes) {
    if (!src || !dst) return fail(BLQH_EINVAL, "null argument");
    return guarded([&] { blqh_device_copy_impl(src, dst, bytes); });
}

int blqh_device_alloc(void** ptr, int64_t bytes) {
    if (!ptr || bytes <= 0) return fail(BLQH_EINVAL, "bad argument");
    return guarded([&] {
        blqh::runtime();
        BLQH_HIP(hipMalloc(ptr, (size_t)bytes));
    });
}

int blqh_device_free(void* ptr) {
    return guarded([&] { if (ptr) BLQH_HIP(hipFree(ptr)); });
}

int blqh_synchronize(void) {
    return guarded([&] { BLQH_HIP(hipStreamSynchronize(blqh::stream())); });
}

int blqh_release_cached(int64_t* bytes) {
    return guarded([&] {
        if (g_rt_ok) BLQH_HIP(hipStreamSynchronize(blqh::stream()));
        const size_t b = blqh::dev_release_cached();
        if (bytes) *bytes = (int64_t)b;
    });
}

int blqh_profile_enable(int on) {
    blqh::profile().on = on != 0;
    return BLQH_OK;
}

int blqh_profile_read(int64_t* launches, double* kernel_ms, double* bytes) {
    const blqh::Profile& p = blqh::profile();
    if (launches) *launches = p.launches;
    if (kernel_ms) *kernel_ms = p.ms;
    if (bytes) *bytes = p.bytes;
    return BLQH_OK;
}

int blqh_profile_reset(void) {
    blqh::Profile& p = blqh::profile();
    p.launches = 0;
    p.ms = 0;
    p.bytes = 0;
    return BLQH_OK;
}

int blqh_start_vector(int64_t n, uint32_t seed, double* out) {
    if (n < 0 || !out) return fail(BLQH_EINVAL, "bad argument");
    return guarded([&] {
        std::vector<double2> v;
        blqh::start_vector(n, seed, v);
        std::memcpy(out, v.data(), n * sizeof(double2));
    });
}

int blqh_tridiag_lowest(int32_t m, const double* alpha, const double* offdiag, double* e0, double* vec) {
    if (m < 1 || !alpha || (m > 1 && !offdiag) || !e0) return fail(BLQH_EINVAL, "bad argument");
    return guarded([&] { *e0 = blqh::tridiag_lowest(m, alpha, offdiag, vec); });
}

}  // extern "C"
