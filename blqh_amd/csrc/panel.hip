// panel.hip -- panel-major copy of the Hamiltonian for the SpMV hot loop.
//
// The CSR rows of a Hamiltonian touch ~350 columns spread over the whole
// vector, so a row-major SpMV gathers x from everywhere: at 2.3M states the
// 37 MB vector falls out of L2 and every 16-byte gather pulls a full line.
// Here the columns are cut into panels of 65536 (x slice of 1 MB for complex
// vectors, which stays resident in each XCD's 4 MB L2).  Panel p stores, for
// every row, that row's entries with p*65536 <= col < (p+1)*65536, rows in
// order, so each panel is one sequential stream:
//   pbase[p]        first entry of panel p
//   prow[p][r]      offset of row r's segment inside panel p (uint32, n+1 per panel)
//   pcol[k]         col - p*65536 (uint16)
//   pvr[k] / pvc[k] value (real double when the matrix is real, else double2)
// The SpMV kernel is persistent: workgroup b owns a contiguous row range and
// keeps those rows' partial sums in LDS while it walks the panels in order, so
// all workgroups sweep the same panel at about the same time and y is written
// once.  The matrix is stored real when every imaginary part is below 1e-14
// of the largest entry (the K = 0 sectors are exactly real; K = nphi/2 sectors
// carry sin(pi n) rounding residue of 1e-16).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "engine.hpp"
#include "panel.hpp"

namespace {

using namespace blqh;

constexpr int kPanelShift = 16;
constexpr int kBlk = 256;

// per row: entries per panel -> cnt[p * n + row]
__global__ void k_panel_count(int64_t n, int npan, const int64_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                              int32_t* __restrict__ cnt) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        int p = 0, c = 0;
        for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) {
            const int q = col[k] >> kPanelShift;
            while (p < q) { cnt[(int64_t)p * n + r] = c; c = 0; ++p; }
            ++c;
        }
        while (p < npan) { cnt[(int64_t)p * n + r] = c; c = 0; ++p; }
    }
}

// off = exclusive scan of cnt over (panel, row): global destination of each segment
__global__ void k_panel_scatter(int64_t n, const int64_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                const double2* __restrict__ val, const int64_t* __restrict__ off, int real,
                                uint16_t* __restrict__ pcol, double* __restrict__ pvr, double2* __restrict__ pvc) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        int p = -1;
        int64_t dst = 0;
        for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) {
            const int q = col[k] >> kPanelShift;
            if (q != p) { p = q; dst = off[(int64_t)q * n + r]; }
            pcol[dst] = (uint16_t)(col[k] & (kPanelWidth - 1));
            if (real) pvr[dst] = val[k].x; else pvc[dst] = val[k];
            ++dst;
        }
    }
}

// prow[p][r] = off[p*n + r] - pbase[p] for r <= n
__global__ void k_panel_rows(int64_t n, int npan, const int64_t* __restrict__ off, int64_t total,
                             uint32_t* __restrict__ prow, int64_t* __restrict__ pbase) {
    const int64_t m = (int64_t)npan * (n + 1);
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < m; t += (int64_t)gridDim.x * blockDim.x) {
        const int p = (int)(t / (n + 1));
        const int64_t r = t - (int64_t)p * (n + 1);
        const int64_t base = off[(int64_t)p * n];
        const int64_t g = (int64_t)p * n + r;  // r == n reads the next panel's first offset (or the total)
        const int64_t v = g < (int64_t)npan * n ? off[g] : total;
        prow[t] = (uint32_t)(v - base);
        if (r == 0) pbase[p] = base;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) pbase[npan] = total;
}

// max |re| and max |im| of the values (two doubles via atomicMax on the bit
// pattern of non-negative doubles, one atomic pair per workgroup)
__global__ void __launch_bounds__(kBlk) k_absmax(int64_t nnz, const double2* __restrict__ val,
                                                 unsigned long long* __restrict__ out) {
    __shared__ double sr[kBlk / 64], si[kBlk / 64];
    double mr = 0, mi = 0;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * blockDim.x) {
        mr = fmax(mr, fabs(val[k].x));
        mi = fmax(mi, fabs(val[k].y));
    }
    for (int o = 32; o > 0; o >>= 1) {
        mr = fmax(mr, __shfl_xor(mr, o, 64));
        mi = fmax(mi, __shfl_xor(mi, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        sr[threadIdx.x >> 6] = mr;
        si[threadIdx.x >> 6] = mi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlk / 64; ++w) {
            mr = fmax(mr, sr[w]);
            mi = fmax(mi, si[w]);
        }
        atomicMax(&out[0], (unsigned long long)__double_as_longlong(mr));
        atomicMax(&out[1], (unsigned long long)__double_as_longlong(mi));
    }
}

// Reorders the entries of every (64-row slice, panel) run between row-major
// order and step-major ("jagged") order: step t holds entry t of each row of
// the slice that has more than t entries in the panel, rows in order.  One
// wavefront per run, lane l <-> row l: at step t the rows still active take
// consecutive positions, off_t + (active lanes below l).  Any number of
// matrices in one launch (grid.y = matrix); each run is written to the
// matrix's scratch arrays, four steps of loads in flight per lane, and
// k_relayout_back copies the scratch over the panel arrays.
__global__ void __launch_bounds__(kBlk) k_relayout_runs(const RelayoutJob* __restrict__ jobs, int to_jagged) {
    const RelayoutJob J = jobs[blockIdx.y];
    const int lane = threadIdx.x & 63;
    const int64_t n = J.n;
    const int64_t nslices = (n + 63) >> 6;
    const int64_t total = nslices * J.npan;
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t w = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); w < total; w += nw) {
        const int p = (int)(w / nslices);
        const int64_t slice = w - (int64_t)p * nslices;
        const uint32_t* rp = J.prow + (int64_t)p * (n + 1);
        const int64_t row0 = slice * 64, myrow = row0 + lane;
        const int64_t rend = row0 + 64 < n ? row0 + 64 : n;
        const uint32_t s = myrow < n ? rp[myrow] : rp[rend];
        const uint32_t e = myrow < n ? rp[myrow + 1] : rp[rend];
        const uint32_t len = e - s;
        const int64_t pb = J.pbase[p];
        const int64_t run = pb + rp[row0];
        int64_t off = 0;
        for (uint32_t t = 0;; t += 4) {
            uint64_t m[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) m[u] = __ballot(t + u < len);
            if (m[0] == 0) break;
            int64_t from[4], to[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int64_t jg = run + off + lanes_below(m[u]);
                const int64_t rm = pb + s + t + u;
                from[u] = to_jagged ? rm : jg;
                to[u] = to_jagged ? jg : rm;
                off += __popcll(m[u]);
            }
            uint16_t c[4] = {0, 0, 0, 0};
            double2 v[4] = {};
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (t + u < len) {
                    c[u] = J.col[from[u]];
                    v[u] = J.real ? make_double2(J.vr[from[u]], 0.0) : J.vc[from[u]];
                }
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (t + u < len) {
                    J.tcol[to[u]] = c[u];
                    if (J.real) J.tvr[to[u]] = v[u].x;
                    else J.tvc[to[u]] = v[u];
                }
        }
    }
}

__global__ void __launch_bounds__(kBlk) k_relayout_back(const RelayoutJob* __restrict__ jobs) {
    const RelayoutJob J = jobs[blockIdx.y];
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < J.nnz; k += (int64_t)gridDim.x * blockDim.x) {
        J.col[k] = J.tcol[k];
        if (J.real) J.vr[k] = J.tvr[k];
        else J.vc[k] = J.tvc[k];
    }
}

}  // namespace

namespace blqh {

namespace {
constexpr size_t kAlign = 256;
size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }
size_t entry_bytes(const blqh_hamil* H) {
    const size_t m = (size_t)std::max<int64_t>(H->nnz, 1);
    return align_up(m * sizeof(uint16_t)) + align_up(m * (H->pan.real ? sizeof(double) : sizeof(double2)));
}
}  // namespace

size_t panel_layout_scratch_bytes(blqh_hamil* const* hs, int count) {
    size_t b = align_up((size_t)count * sizeof(RelayoutJob));
    for (int i = 0; i < count; ++i) b += entry_bytes(hs[i]);
    return b;
}

void set_panel_layouts(blqh_hamil* const* hs, int count, bool jagged, void* scratch, size_t scratch_bytes) {
    std::vector<blqh_hamil*> todo;
    for (int i = 0; i < count; ++i)
        if (hs[i]->pan.built && hs[i]->pan.jagged != jagged) todo.push_back(hs[i]);
    if (todo.empty()) return;
    BLQH_REQUIRE(todo.size() <= 65535, BLQH_ERANGE, "too many matrices for one panel relayout");
    hipStream_t st = stream();
    const size_t need = panel_layout_scratch_bytes(todo.data(), (int)todo.size());
    DevBuf<char> own;
    char* sc = (char*)scratch;
    if (!sc || scratch_bytes < need) {
        own.alloc(need);
        sc = own.get();
    }
    std::vector<RelayoutJob> jobs(todo.size());
    size_t off = align_up(todo.size() * sizeof(RelayoutJob));
    int64_t max_runs = 1, max_nnz = 1;
    for (size_t i = 0; i < todo.size(); ++i) {
        blqh_hamil* H = todo[i];
        RelayoutJob& J = jobs[i];
        J.n = H->n;
        J.npan = H->pan.npan;
        J.real = H->pan.real ? 1 : 0;
        J.nnz = H->nnz;
        J.pbase = H->pan.pbase.get();
        J.prow = H->pan.prow.get();
        J.col = H->pan.pcol.get();
        J.vr = H->pan.pvr.get();
        J.vc = H->pan.pvc.get();
        const size_t m = (size_t)std::max<int64_t>(H->nnz, 1);
        J.tcol = (uint16_t*)(sc + off);
        off += align_up(m * sizeof(uint16_t));
        J.tvr = J.real ? (double*)(sc + off) : nullptr;
        J.tvc = J.real ? nullptr : (double2*)(sc + off);
        off += align_up(m * (J.real ? sizeof(double) : sizeof(double2)));
        max_runs = std::max<int64_t>(max_runs, ceil_div(H->n, 64) * H->pan.npan);
        max_nnz = std::max<int64_t>(max_nnz, H->nnz);
    }
    // earlier work on the library stream may still read this scratch; then a
    // synchronous copy of the job table (the host vector goes out of scope)
    BLQH_HIP(hipStreamSynchronize(st));
    BLQH_HIP(hipMemcpy(sc, jobs.data(), jobs.size() * sizeof(RelayoutJob), hipMemcpyHostToDevice));
    const RelayoutJob* dj = (const RelayoutJob*)sc;
    const int gx = (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(max_runs, kBlk / 64), 4096));
    hipLaunchKernelGGL(k_relayout_runs, dim3(gx, (unsigned)jobs.size()), dim3(kBlk), 0, st, dj, jagged ? 1 : 0);
    const int gb = (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(max_nnz, kBlk), 4096));
    hipLaunchKernelGGL(k_relayout_back, dim3(gb, (unsigned)jobs.size()), dim3(kBlk), 0, st, dj);
    BLQH_HIP(hipGetLastError());
    if (own.get()) BLQH_HIP(hipStreamSynchronize(st));  // the temporary is freed on return
    for (auto* H : todo) H->pan.jagged = jagged;
}

void set_panel_layout(blqh_hamil* H, bool jagged, void* scratch, size_t scratch_bytes) {
    set_panel_layouts(&H, 1, jagged, scratch, scratch_bytes);
}

size_t panel_layout_scratch_bytes(const blqh_hamil* H) {
    blqh_hamil* h = const_cast<blqh_hamil*>(H);
    return panel_layout_scratch_bytes(&h, 1);
}

void build_panels(blqh_hamil* H) {
    hipStream_t st = stream();
    const int64_t n = H->n;
    const int npan = (int)std::max<int64_t>(1, ceil_div(n, (int64_t)1 << kPanelShift));
    Panels& P = H->pan;
    P.npan = npan;
    DevBuf<unsigned long long> mx(2);
    BLQH_HIP(hipMemsetAsync(mx.get(), 0, 16, st));
    const int g = (int)std::min<int64_t>(std::max<int64_t>(ceil_div(H->nnz, kBlk * 8), 1), 2048);
    hipLaunchKernelGGL(k_absmax, dim3(g), dim3(kBlk), 0, st, H->nnz, H->val.get(), mx.get());
    unsigned long long hm[2];
    BLQH_HIP(hipMemcpyAsync(hm, mx.get(), 16, hipMemcpyDeviceToHost, st));
    BLQH_HIP(hipStreamSynchronize(st));
    double mre, mim;
    std::memcpy(&mre, &hm[0], 8);
    std::memcpy(&mim, &hm[1], 8);
    P.real = mim <= 1e-14 * std::max(mre, mim);
    DevBuf<int32_t> cnt((size_t)npan * n);
    DevBuf<int64_t> off((size_t)npan * n + 1);
    const int gr = (int)std::min<int64_t>(std::max<int64_t>(ceil_div(n, kBlk), 1), 65536);
    hipLaunchKernelGGL(k_panel_count, dim3(gr), dim3(kBlk), 0, st, n, npan, H->rowptr.get(), H->col.get(), cnt.get());
    BLQH_HIP(hipGetLastError());
    exclusive_scan(cnt.get(), off.get(), (int64_t)npan * n, st);
    P.prow.alloc((size_t)npan * (n + 1));
    P.pbase.alloc(npan + 1);
    const int gp = (int)std::min<int64_t>(ceil_div((int64_t)npan * (n + 1), kBlk), 65536);
    hipLaunchKernelGGL(k_panel_rows, dim3(gp), dim3(kBlk), 0, st, n, npan, off.get(), H->nnz, P.prow.get(),
                       P.pbase.get());
    P.pcol.alloc(std::max<int64_t>(H->nnz, 1));
    if (P.real) P.pvr.alloc(std::max<int64_t>(H->nnz, 1));
    else P.pvc.alloc(std::max<int64_t>(H->nnz, 1));
    hipLaunchKernelGGL(k_panel_scatter, dim3(gr), dim3(kBlk), 0, st, n, H->rowptr.get(), H->col.get(), H->val.get(),
                       off.get(), P.real ? 1 : 0, P.pcol.get(), P.pvr.get(), P.pvc.get());
    BLQH_HIP(hipGetLastError());
    BLQH_HIP(hipStreamSynchronize(st));  // cnt / off are freed on return
    P.built = true;
}

}  // namespace blqh
