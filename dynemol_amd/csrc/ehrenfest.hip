// Diabatic-Ehrenfest kernel matrix K = X o A - H' A on the device, from the
// propagator's resident state (diabatic-Ehren.f:100-116 and :366-392,
// Taylor_gpu.cpp:800-875, Chebyshev_gpu_kernels.cu:358-455).
//
//   rho_ij = Re(AO_ket_j,e AO_bra_i,e) - Re(AO_ket_j,h AO_bra_i,h)
//   A      = (rho + rho^T) / 2
//   K      = X o A - H' A,   X_ij the extended-Hueckel factors (h_ij = X_ij S_ij)
//
// One elementwise pass writes A and X o A (the latter into K); the N^3 product
// is the library's FP64 matrix-core GEMM with C = X o A in the accumulators
// (alpha = -1, beta = 1).
#include "kernels.hpp"

namespace dh {

// Re(a b) written as in diabatic-Ehren.f:364
__device__ inline double re_mul(cplx a, cplx b) { return a.re * b.re - a.im * b.im; }

// Thread (i, j): i along the 256 threads of a block (coalesced column
// segments of A and K), j = blockIdx.y.  Rows and columns >= N of A and K are
// written as 0.
__global__ __launch_bounds__(256) void ehrenfest_prep_kernel(EhrenfestArgs e) {
    const int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
    if (i >= e.ld) return;
    const long o = (long)j * e.ld + i;
    if (i >= e.N || j >= e.N) {
        e.A[o] = 0.0;
        e.K[o] = 0.0;
        return;
    }
    auto bra = [&](int r, int col) {
        cplx v = e.bra[(long)r * e.bra_stride + col];
        if (e.conj_bra) v.im = -v.im;
        return v;
    };
    auto ket = [&](int r, int col) { return e.ket[(long)r * e.ket_stride + col]; };
    double rij = re_mul(ket(j, e.ket_el), bra(i, e.bra_el));
    double rji = re_mul(ket(i, e.ket_el), bra(j, e.bra_el));
    if (e.bra_hl >= 0) {
        rij -= re_mul(ket(j, e.ket_hl), bra(i, e.bra_hl));
        rji -= re_mul(ket(i, e.ket_hl), bra(j, e.bra_hl));
    }
    const double a = (rij + rji) * 0.5;
    double x;
    if (e.X) x = e.X[(long)j * e.ldx + i];
    else if (i == j) x = e.ip[i] + e.vshift[i];
    else x = huckel_x(e.ip[i], e.ip[j], e.kwh[i], e.kwh[j], e.vshift[i], e.vshift[j]);
    e.A[o] = a;
    e.K[o] = x * a;
}

// dst = the lower triangle of src (diagonal included), zero above
__global__ __launch_bounds__(256) void copy_lower_kernel(const double* __restrict__ src, double* __restrict__ dst,
                                                         int ld) {
    const int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
    if (i >= ld) return;
    const long o = (long)j * ld + i;
    dst[o] = i >= j ? src[o] : 0.0;
}

void launch_copy_lower(const double* src, double* dst, int ld, hipStream_t st) {
    hipLaunchKernelGGL(copy_lower_kernel, dim3((ld + 255) / 256, ld), dim3(256), 0, st, src, dst, ld);
}

void launch_ehrenfest(const EhrenfestArgs& e, const double* Hp, hipStream_t st) {
    hipLaunchKernelGGL(ehrenfest_prep_kernel, dim3((e.ld + 255) / 256, e.ld), dim3(256), 0, st, e);
    Gemm g{};
    g.M = e.N; g.N = e.N; g.K = e.N;
    g.A = Hp; g.lda = e.ld;
    g.B = e.A; g.ldb = e.ld;
    g.C = e.K; g.ldc = e.ld;
    g.alpha = -1.0; g.beta = 1.0; g.kmode = K_FULL; g.batch = 1;
    launch_gemm(g, st);
}

}  // namespace dh
