/*
 * Test hooks of libdynemol_hip: run one internal kernel on host-supplied
 * operands (copied to and from the device).  For the library's own tests;
 * no counterpart in the reference.
 */
#ifndef DYNEMOL_HIP_TESTING_H
#define DYNEMOL_HIP_TESTING_H
#ifdef __cplusplus
extern "C" {
#endif
/* C = alpha op(A) op(B) + beta C through the FP64 MFMA GEMM kernel; kmode and
 * lower_only as the internal k-range / lower-tile options (0 = full). */
int dh_test_gemm(int M, int N, int K, const double *A, int lda, const double *B, int ldb, double *C, int ldc,
                 double alpha, double beta, int transA, int transB, int kmode, int lower_only);
/* In-place blocked Cholesky of an n x n SPD matrix (n a multiple of 64),
 * returning L (lower, upper zeroed) and L^-1 (lower) and the device info code. */
int dh_test_cholesky(int n, double *A, double *Linv, int *info);
#ifdef __cplusplus
}
#endif
#endif
