// tiles.hip -- column-tile copy of the Hamiltonian for SpMV kind 4.
//
// Layout (panel.hpp, Tiles): G row blocks of SB 64-row slices, one per
// workgroup of 16 wavefronts; wave w of block b owns slices b*SB + w + 16 r,
// r < R, so lane l of the wave owns R rows ("slots" r).  The columns are cut
// into tiles of kTileCols.  Each wave's entries form one contiguous stream:
// for tile t = 0.., for step j = 0.., for slot r = 0..R-1, entry j of every
// lane's row in slot r that has more than j entries in the tile (lanes in
// order).  At one step a wave therefore loads the entries of all R slots at
// once, which keeps enough bytes in flight to stream HBM.  The SpMV stages
// the x tile in LDS, so every gather is an LDS read.
//
// Build from the CSR (columns ascending per row):
//   k_tile_len     thread per row: entries per (row, tile) -> len (u8), overflow flag
//   k_tile_wavelen thread per wave stream: entries of the stream; exclusive scan -> wstart
//   k_tile_scatter wave per wave stream: every tile written in step-major order
#include <algorithm>

#include "engine.hpp"

namespace {

using namespace blqh;

constexpr int kBlk = 256;

struct TileGeom {
    int64_t n, nslices;
    int G, SB, R, ntiles;
};

// index of (row's wave stream, tile t, slice r of the wave, lane) in len
__device__ __forceinline__ int64_t len_index(const TileGeom& g, int64_t row, int t) {
    const int64_t s = row >> 6;
    const int lane = (int)(row & 63);
    const int64_t b = s / g.SB, sr = s - b * g.SB;
    const int64_t w = sr % kTileWaves, r = sr / kTileWaves;
    return (((b * kTileWaves + w) * g.ntiles + t) * g.R + r) * 64 + lane;
}

__global__ void __launch_bounds__(kBlk) k_tile_len(TileGeom g, const int64_t* __restrict__ rowptr,
                                                   const int32_t* __restrict__ col, uint8_t* __restrict__ len,
                                                   int* __restrict__ overflow) {
    for (int64_t row = blockIdx.x * (int64_t)kBlk + threadIdx.x; row < g.n; row += (int64_t)gridDim.x * kBlk) {
        int cur = -1, c = 0;
        for (int64_t k = rowptr[row]; k < rowptr[row + 1]; ++k) {
            const int t = col[k] / kTileCols;
            if (t != cur) {
                if (cur >= 0) {
                    if (c > 255) atomicOr(overflow, 1);
                    len[len_index(g, row, cur)] = (uint8_t)(c > 255 ? 255 : c);
                }
                cur = t;
                c = 0;
            }
            ++c;
        }
        if (cur >= 0) {
            if (c > 255) atomicOr(overflow, 1);
            len[len_index(g, row, cur)] = (uint8_t)(c > 255 ? 255 : c);
        }
    }
}

// entries of each wave stream (sum of its len bytes)
__global__ void __launch_bounds__(kBlk) k_tile_wavelen(int64_t nwaves, int64_t bytes_per_wave,
                                                       const uint8_t* __restrict__ len, int32_t* __restrict__ wl) {
    const int lane = threadIdx.x & 63;
    const int64_t nw = (int64_t)gridDim.x * (kBlk / 64);
    for (int64_t w = blockIdx.x * (int64_t)(kBlk / 64) + (threadIdx.x >> 6); w < nwaves; w += nw) {
        const uint8_t* p = len + w * bytes_per_wave;
        int s = 0;
        for (int64_t i = lane; i < bytes_per_wave; i += 64) s += p[i];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        if (lane == 0) wl[w] = s;
    }
}

// One wavefront per wave stream: lane l walks the CSR entries of its R rows
// (ascending columns, hence ascending tiles) and writes each tile step-major.
template <bool REAL>
__global__ void __launch_bounds__(kBlk) k_tile_scatter(TileGeom g, const int64_t* __restrict__ rowptr,
                                                       const int32_t* __restrict__ col,
                                                       const double2* __restrict__ val,
                                                       const uint8_t* __restrict__ len,
                                                       const int64_t* __restrict__ wstart,
                                                       uint16_t* __restrict__ tcol, double* __restrict__ tvr,
                                                       double2* __restrict__ tvc) {
    const int lane = threadIdx.x & 63;
    const int64_t nwaves = (int64_t)g.G * kTileWaves;
    const int64_t nw = (int64_t)gridDim.x * (kBlk / 64);
    for (int64_t wid = blockIdx.x * (int64_t)(kBlk / 64) + (threadIdx.x >> 6); wid < nwaves; wid += nw) {
        const int64_t b = wid / kTileWaves, w = wid % kTileWaves;
        int64_t cur[kTileMaxR];
        for (int r = 0; r < g.R; ++r) {
            const int64_t slice = b * g.SB + w + (int64_t)kTileWaves * r;
            const int64_t row = slice * 64 + lane;
            cur[r] = slice < (b + 1) * g.SB && row < g.n ? rowptr[row] : 0;
        }
        int64_t ptr = wstart[wid];
        const uint8_t* L = len + wid * (int64_t)g.ntiles * g.R * 64 + lane;
        for (int t = 0; t < g.ntiles; ++t) {
            uint32_t l[kTileMaxR];
            uint32_t lmax = 0;
            for (int r = 0; r < g.R; ++r) {
                l[r] = L[((int64_t)t * g.R + r) * 64];  // zero for rows beyond the matrix
                lmax = l[r] > lmax ? l[r] : lmax;
            }
            int64_t off = 0;
            for (uint32_t j = 0; __ballot(j < lmax) != 0; ++j) {
                for (int r = 0; r < g.R; ++r) {
                    const uint64_t m = __ballot(j < l[r]);
                    if (j < l[r]) {
                        const int64_t pos = ptr + off + lanes_below(m);
                        const int64_t k = cur[r] + j;
                        tcol[pos] = (uint16_t)(col[k] - t * kTileCols);
                        if (REAL) tvr[pos] = val[k].x;
                        else tvc[pos] = val[k];
                    }
                    off += __popcll(m);
                }
            }
            ptr += off;
            for (int r = 0; r < g.R; ++r) cur[r] += l[r];
        }
    }
}

}  // namespace

namespace blqh {

bool build_tiles(blqh_hamil* H, int G) {
    Tiles& T = H->til;
    if (T.built) return true;
    if (T.tried) return false;
    T.tried = true;
    const int64_t n = H->n;
    if (n < 64 || !H->pan.built) return false;
    TileGeom g;
    g.n = n;
    g.nslices = ceil_div(n, 64);
    g.G = G;
    g.SB = (int)ceil_div(g.nslices, G);
    g.R = (int)ceil_div(g.SB, kTileWaves);
    g.ntiles = (int)ceil_div(n, kTileCols);
    if (g.R > kTileMaxR) return false;
    hipStream_t st = stream();
    const int64_t nwaves = (int64_t)G * kTileWaves;
    const int64_t nruns = nwaves * g.ntiles * g.R;
    if (nruns >= (int64_t)1 << 31 || H->nnz >= (int64_t)1 << 40) return false;
    DevBuf<uint8_t> len((size_t)nruns * 64);
    DevBuf<int> ovf(1);
    BLQH_HIP(hipMemsetAsync(len.get(), 0, len.bytes(), st));
    BLQH_HIP(hipMemsetAsync(ovf.get(), 0, sizeof(int), st));
    const int gr = (int)std::min<int64_t>(std::max<int64_t>(ceil_div(n, kBlk), 1), 65536);
    const CsrView A = spmv_csr(H);
    hipLaunchKernelGGL(k_tile_len, dim3(gr), dim3(kBlk), 0, st, g, A.rowptr, A.col, len.get(),
                       ovf.get());
    int hovf = 0;
    BLQH_HIP(hipMemcpyAsync(&hovf, ovf.get(), sizeof(int), hipMemcpyDeviceToHost, st));
    BLQH_HIP(hipStreamSynchronize(st));
    if (hovf) return false;  // a row has more than 255 entries in one tile
    DevBuf<int32_t> wl(nwaves);
    DevBuf<int64_t> ws(nwaves + 1);
    const int gw = (int)std::min<int64_t>(std::max<int64_t>(ceil_div(nwaves, kBlk / 64), 1), 65536);
    hipLaunchKernelGGL(k_tile_wavelen, dim3(gw), dim3(kBlk), 0, st, nwaves, (int64_t)g.ntiles * g.R * 64, len.get(),
                       wl.get());
    exclusive_scan(wl.get(), ws.get(), nwaves, st);
    const int64_t m = std::max<int64_t>(H->nnz, 1);
    T.real = H->pan.real;
    T.col.alloc(m);
    if (T.real) T.vr.alloc(m);
    else T.vc.alloc(m);
    if (T.real)
        hipLaunchKernelGGL(k_tile_scatter<true>, dim3(gw), dim3(kBlk), 0, st, g, A.rowptr, A.col,
                           A.val, len.get(), ws.get(), T.col.get(), T.vr.get(), T.vc.get());
    else
        hipLaunchKernelGGL(k_tile_scatter<false>, dim3(gw), dim3(kBlk), 0, st, g, A.rowptr, A.col,
                           A.val, len.get(), ws.get(), T.col.get(), T.vr.get(), T.vc.get());
    BLQH_HIP(hipGetLastError());
    BLQH_HIP(hipStreamSynchronize(st));  // wl is freed on return
    T.len = std::move(len);
    T.wstart = std::move(ws);
    T.G = g.G;
    T.SB = g.SB;
    T.R = g.R;
    T.ntiles = g.ntiles;
    T.built = true;
    return true;
}

}  // namespace blqh
