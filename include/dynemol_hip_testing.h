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
/* A plain device copy of `bytes` (read + write), repeated; for calibrating
 * memory-traffic counters against a kernel of known traffic. */
int dh_test_copy(long long bytes, int repeats, double *ms_per_copy);
/* A plain streaming read of `bytes` (16-byte loads, 2048 workgroups), repeated:
 * the read bandwidth a kernel streaming an operator of that size can reach. */
int dh_test_read(long long bytes, int repeats, double *ms_per_read);
/* Average launch time (microseconds, HIP events) of one matvec variant on an
 * N x N FP64 operator with R complex right-hand sides (Chebyshev epilogue);
 * variant -1 times the FP32 matrix-core GEMM of the mixed mode, -2 the term
 * kernel of the FP32-stored operator mode. */
int dh_test_matvec_timing(int N, int R, int variant, int reps, double *us_per_launch);
/* Force every matvec to one kernel variant (-1 restores the autotuned choice). */
int dh_test_force_matvec_variant(int variant);
/* Device memory for tests of the entry points that take device pointers:
 * allocate, free, and copy to (to_device = 1) or from (0) the device. */
int dh_test_device_alloc(long long bytes, void **ptr);
int dh_test_device_free(void *ptr);
int dh_test_memcpy(void *dst, const void *src, long long bytes, int to_device);
/* The Chebyshev coefficients of exp(-i x T) e^{-i phase} as the propagation
 * computes them (host code, no device work): coef (2 (kmax + 2) doubles,
 * complex pairs) = (2 - delta_k0) (-i)^k J_k(x) e^{-i phase} for k = 0..K and
 * 0 at K + 1, where K is one past the last k with |J_k(x)| > tol (at most the
 * recurrence's start index), and *K_out = K. */
int dh_test_bessel_coefficients(double x, double phase, double tol, int kmax, double *coef, int *K_out);
/* Set the batch term family used for R right-hand sides (1 symmetric half
 * storage, 0 dense column blocks; -1: the autotuner chooses again), for the
 * batches created afterwards. */
int dh_test_force_batch_family(int R, int family);
#ifdef __cplusplus
}
#endif
#endif
