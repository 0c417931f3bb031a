// Library calls of the build, through one rocBLAS handle per context:
// * generalised eigenpairs h C = S C E (C^T S C = I) of a system's resident
//   operators, for initial states: the reference's FMO step solves the same
//   pencil with LAPACK/MAGMA dsygvd (FMO.f:294-350, GPU_Interface.cpp:705-853);
// * the plain full FP64 GEMM of the Ehrenfest kernel matrix (rocBLAS dgemm ran
//   the 4096^3 product in 2.2 ms on this part, the library's own MFMA kernel
//   in 2.8 ms; the rebuild's triangular products stay on the library's
//   kernels, which skip the zero half).
//
// H_orth = L^-1 h L^-T is symmetric with the pencil's eigenvalues, so the
// eigenpairs (ascending) come from rocSOLVER's dsyevd on the device, and
// C = L^-T V from the library's own triangular FP64 GEMM.
// Once per request, off the propagation path.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "../../include/dynemol_hip.h"
#include "kernels.hpp"

namespace dh {

struct LibHandle {
    rocblas_handle h = nullptr;
};

LibHandle* lib_create(hipStream_t st) {
    auto* es = new LibHandle;
    if (rocblas_create_handle(&es->h) != rocblas_status_success) {
        delete es;
        return nullptr;
    }
    rocblas_set_stream(es->h, st);
    // kernels without atomics: the same bits on every run, which the term
    // products of the mixed modes rely on (a request alone and in a mix)
    rocblas_set_atomics_mode(es->h, rocblas_atomics_not_allowed);
    return es;
}

void lib_destroy(LibHandle* es) {
    if (!es) return;
    if (es->h) rocblas_destroy_handle(es->h);
    delete es;
}

int lib_dgemm_nn(LibHandle* es, int m, int n, int k, double alpha, const double* A, int lda, const double* B, int ldb,
                 double beta, double* C, int ldc) {
    const rocblas_status rs = rocblas_dgemm(es->h, rocblas_operation_none, rocblas_operation_none, m, n, k, &alpha, A,
                                            lda, B, ldb, &beta, C, ldc);
    if (rs != rocblas_status_success) {
        set_error("rocblas_dgemm returned status %d", (int)rs);
        return DH_ERR_HIP;
    }
    return DH_OK;
}

// The mixed modes' term products (FP32 and bf16 inputs, FP32 accumulation):
// strided-batched over K slices; rocBLAS ran these at 0.87-0.95 of the FP32
// matrix peak for N = 4096-8192 with 128-256 real columns in 4 slices, where the
// library's own split-K FP32 GEMM reached 0.37.
int lib_gemm_slices_f32(LibHandle* es, int m, int c, int KS, int S, const float* A, int lda, const float* X, int ldx,
                        float* Y, int ldy) {
    const float one = 1.f, zero = 0.f;
    const rocblas_status rs = rocblas_sgemm_strided_batched(
        es->h, rocblas_operation_none, rocblas_operation_transpose, m, c, KS, &one, A, lda, (rocblas_stride)KS * lda,
        X, ldx, (rocblas_stride)KS * ldx, &zero, Y, ldy, (rocblas_stride)c * ldy, S);
    if (rs != rocblas_status_success) {
        set_error("rocblas_sgemm_strided_batched returned status %d", (int)rs);
        return DH_ERR_HIP;
    }
    return DH_OK;
}

int lib_gemm_slices_bf16(LibHandle* es, int m, int c, int KS, int S, const unsigned short* A, int lda,
                         const unsigned short* X, int ldx, float* Y, int ldy) {
    const float one = 1.f, zero = 0.f;
    const rocblas_status rs = rocblas_gemm_strided_batched_ex(
        es->h, rocblas_operation_none, rocblas_operation_transpose, m, c, KS, &one, A, rocblas_datatype_bf16_r, lda,
        (rocblas_stride)KS * lda, X, rocblas_datatype_bf16_r, ldx, (rocblas_stride)KS * ldx, &zero, Y,
        rocblas_datatype_f32_r, ldy, (rocblas_stride)c * ldy, Y, rocblas_datatype_f32_r, ldy, (rocblas_stride)c * ldy,
        S, rocblas_datatype_f32_r, rocblas_gemm_algo_standard, 0, 0);
    if (rs != rocblas_status_success) {
        set_error("rocblas_gemm_strided_batched_ex (bf16) returned status %d", (int)rs);
        return DH_ERR_HIP;
    }
    return DH_OK;
}

// All eigenpairs of H_orth by divide and conquer (dsyevd: 0.15 s at n = 4096
// on this part, against 3-4 s for the index-range solvers dsyevx/dsyevdx),
// then C = L^-T V for the selected columns only.
// A, Z: ld x ld scratch; C: ld x (iu - il + 1); E: 2 N doubles on the device
// (the eigenvalues, then the tridiagonal's off-diagonal workspace); the
// selected eigenvalues are E[il - 1 .. iu - 1].  Q2 holds L^-1 in its lower
// triangle (its upper triangle is not read).  Returns 0, or sets the error
// text and returns a DH_ERR code.
int eigen_pairs(LibHandle* es, int N, int ld, const double* Hth, const double* Q2, double* A, double* Z, double* C,
                double* E, int il, int iu, int* d_ints, hipStream_t st) {
    const int nev = iu - il + 1;
    launch_copy_matrix(Hth, A, (long)ld * ld, st);
    rocblas_status rs = rocsolver_dsyevd(es->h, rocblas_evect_original, rocblas_fill_lower, N, A, ld, E, E + N, d_ints);
    if (rs != rocblas_status_success) {
        set_error("eigenpairs: rocsolver_dsyevd returned status %d", (int)rs);
        return DH_ERR_HIP;
    }
    int info = 0;
    DH_HIP(hipMemcpyAsync(&info, d_ints, sizeof(int), hipMemcpyDeviceToHost, st));
    DH_HIP(hipStreamSynchronize(st));
    if (info != 0) {
        set_error("eigenpairs: dsyevd did not converge (info %d)", info);
        return DH_ERR_DIVERGED;
    }
    // C = L^-T V(:, il..iu): L^-1 with a zero upper triangle, k from the row tile on
    launch_copy_lower(Q2, Z, ld, st);
    Gemm g{};
    g.M = N; g.N = nev; g.K = N;
    g.A = Z; g.lda = ld; g.transA = 1;
    g.B = A + (long)ld * (il - 1); g.ldb = ld;
    g.C = C; g.ldc = ld;
    g.alpha = 1.0; g.beta = 0.0; g.kmode = K_FROM_ROWTILE; g.batch = 1;
    launch_gemm(g, st);
    DH_HIP(hipGetLastError());
    return DH_OK;
}

}  // namespace dh
