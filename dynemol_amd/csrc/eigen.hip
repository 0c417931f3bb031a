// Generalised eigenpairs h C = S C E (C^T S C = I) of a system's resident
// operators, for initial states: the reference's FMO step solves the same
// pencil with LAPACK/MAGMA dsygvd (FMO.f:294-350, GPU_Interface.cpp:705-853).
//
// H_orth = L^-1 h L^-T is symmetric with the pencil's eigenvalues, so the
// selected eigenpairs (by index, ascending) come from rocSOLVER's dsyevdx on
// the device, and C = L^-T V from the library's own triangular FP64 GEMM.
// Once per request, off the propagation path.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "../../include/dynemol_hip.h"
#include "kernels.hpp"

namespace dh {

struct EigenSolver {
    rocblas_handle h = nullptr;
};

EigenSolver* eigen_create(hipStream_t st) {
    auto* es = new EigenSolver;
    if (rocblas_create_handle(&es->h) != rocblas_status_success) {
        delete es;
        return nullptr;
    }
    rocblas_set_stream(es->h, st);
    return es;
}

void eigen_destroy(EigenSolver* es) {
    if (!es) return;
    if (es->h) rocblas_destroy_handle(es->h);
    delete es;
}

// A: ld x ld scratch (destroyed); Z: ld x N (dsyevdx may use every column as
// workspace); C: ld x (iu - il + 1); E: N doubles on the
// device (the first iu - il + 1 hold the eigenvalues).  Q2 holds L^-1 in its
// lower triangle (its upper triangle is not read).  Returns 0, or sets the
// error text and returns a DH_ERR code.
int eigen_pairs(EigenSolver* es, int N, int ld, const double* Hth, const double* Q2, double* A, double* Z, double* C,
                double* E, int il, int iu, int* d_ints, hipStream_t st) {
    const int nev = iu - il + 1;
    launch_copy_matrix(Hth, A, (long)ld * ld, st);
    rocblas_status rs = rocsolver_dsyevdx(es->h, rocblas_evect_original, rocblas_erange_index, rocblas_fill_lower, N,
                                          A, ld, 0.0, 0.0, il, iu, d_ints, E, Z, ld, d_ints + 1);
    if (rs != rocblas_status_success) {
        set_error("eigenpairs: rocsolver_dsyevdx returned status %d", (int)rs);
        return DH_ERR_HIP;
    }
    int h_ints[2] = {0, 0};
    DH_HIP(hipMemcpyAsync(h_ints, d_ints, sizeof(int) * 2, hipMemcpyDeviceToHost, st));
    DH_HIP(hipStreamSynchronize(st));
    if (h_ints[1] != 0 || h_ints[0] != nev) {
        set_error("eigenpairs: dsyevdx did not converge (info %d, %d of %d eigenpairs)", h_ints[1], h_ints[0], nev);
        return DH_ERR_DIVERGED;
    }
    // C = L^-T Z: L^-1 with a zero upper triangle (A is free again), k from the row tile on
    launch_copy_lower(Q2, A, ld, st);
    Gemm g{};
    g.M = N; g.N = nev; g.K = N;
    g.A = A; g.lda = ld; g.transA = 1;
    g.B = Z; g.ldb = ld;
    g.C = C; g.ldc = ld;
    g.alpha = 1.0; g.beta = 0.0; g.kmode = K_FROM_ROWTILE; g.batch = 1;
    launch_gemm(g, st);
    DH_HIP(hipGetLastError());
    return DH_OK;
}

}  // namespace dh
