// dense.hip -- dense diagonalisation cross-check on the GPU (rocSOLVER zheevd).
//
// Reference: hamil::diag + hamil::ground_state_energy,
// BLQH_without_spin/hamiltonian.cpp:341-353,413-432 (spinful:
// BLQH_with_spin/hamiltonian.cpp:1131-1144).  The reference fills a row-major
// N x N array with H and hands it to zheevd, which reads it column-major, i.e.
// it diagonalises H^T; the printed energy is Re<H psi_0|psi_0> for the first
// eigenvector psi_0 of H^T.  Both quantities are returned: e0 (lowest
// eigenvalue) and e_printed.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <vector>

#include "engine.hpp"

namespace {
__global__ void k_csr_to_dense(int64_t n, const int64_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                               const double2* __restrict__ val, double2* __restrict__ A) {
    for (int64_t r = blockIdx.x; r < n; r += gridDim.x)
        for (int64_t k = rowptr[r] + threadIdx.x; k < rowptr[r + 1]; k += blockDim.x) A[r * n + col[k]] = val[k];
}
}  // namespace

void blqh_dense_ground_state_impl(const blqh_hamil* H, double* e0, double* e_printed, double* vec) {
    using namespace blqh;
    const int64_t n = H->n;
    BLQH_REQUIRE(n <= 16384, BLQH_ERANGE, "dense diagonalisation limited to 16384 states");
    hipStream_t st = stream();
    DevBuf<double2> A((size_t)n * n);
    DevBuf<double> D(n), E(n);
    DevBuf<int> info(1);
    BLQH_HIP(hipMemsetAsync(A.get(), 0, (size_t)n * n * sizeof(double2), st));
    hipLaunchKernelGGL(k_csr_to_dense, dim3((unsigned)std::min<int64_t>(n, 65535)), dim3(64), 0, st, n,
                       H->rowptr.get(), H->col.get(), H->val.get(), A.get());
    BLQH_HIP(hipGetLastError());
    // the handle is destroyed on every exit, including a throwing copy below
    struct Handle {
        rocblas_handle h = nullptr;
        ~Handle() {
            if (h) rocblas_destroy_handle(h);
        }
    } hd_;
    BLQH_REQUIRE(rocblas_create_handle(&hd_.h) == rocblas_status_success, BLQH_EDEVICE,
                 "rocblas_create_handle failed");
    rocblas_handle h = hd_.h;
    rocblas_set_stream(h, st);
    rocblas_status s = rocsolver_zheevd(h, rocblas_evect_original, rocblas_fill_upper, (rocblas_int)n,
                                        (rocblas_double_complex*)A.get(), (rocblas_int)n, D.get(), E.get(), info.get());
    std::vector<double> hd(n);
    std::vector<double2> v0(n);
    int hinfo = 0;
    BLQH_HIP(hipMemcpyAsync(hd.data(), D.get(), n * 8, hipMemcpyDeviceToHost, st));
    BLQH_HIP(hipMemcpyAsync(v0.data(), A.get(), n * 16, hipMemcpyDeviceToHost, st));
    BLQH_HIP(hipMemcpyAsync(&hinfo, info.get(), sizeof(int), hipMemcpyDeviceToHost, st));
    BLQH_HIP(hipStreamSynchronize(st));
    BLQH_REQUIRE(s == rocblas_status_success && hinfo == 0, BLQH_EINTERNAL, "rocsolver_zheevd failed");
    // Re<H psi_0 | psi_0> with the CSR on the host
    std::vector<int64_t> rp(n + 1);
    std::vector<int32_t> cl(H->nnz);
    std::vector<double2> vl(H->nnz);
    copy_sync(rp.data(), H->rowptr.get(), (n + 1) * 8, hipMemcpyDeviceToHost);
    copy_sync(cl.data(), H->col.get(), H->nnz * 4, hipMemcpyDeviceToHost);
    copy_sync(vl.data(), H->val.get(), H->nnz * 16, hipMemcpyDeviceToHost);
    double e = 0;
    for (int64_t i = 0; i < n; ++i) {
        double yr = 0, yi = 0;
        for (int64_t k = rp[i]; k < rp[i + 1]; ++k) {
            const double2 a = vl[k], x = v0[cl[k]];
            yr += a.x * x.x - a.y * x.y;
            yi += a.x * x.y + a.y * x.x;
        }
        e += yr * v0[i].x + yi * v0[i].y;  // Re(conj(y_i) psi_i)
    }
    if (e0) *e0 = hd[0];
    if (e_printed) *e_printed = e;
    if (vec) std::copy((const double*)v0.data(), (const double*)v0.data() + 2 * n, vec);
}
