// Diabatic-Ehrenfest kernel matrix K = X o A - H' A on the device, from the
// propagator's resident state (diabatic-Ehren.f:100-116 and :366-392,
// Taylor_gpu.cpp:800-875, Chebyshev_gpu_kernels.cu:358-455).
//
//   rho_ij = Re(AO_ket_j,e AO_bra_i,e) - Re(AO_ket_j,h AO_bra_i,h)
//   A      = (rho + rho^T) / 2
//   K      = X o A - H' A,   X_ij the extended-Hueckel factors (h_ij = X_ij S_ij)
//
// One elementwise pass writes A and X o A (the latter into K); the N^3 product
// K <- K - H' A is rocBLAS dgemm (alpha = -1, beta = 1), or the library's FP64
// matrix-core GEMM with C = X o A in the accumulators when no handle is given.
#include "kernels.hpp"

namespace dh {

// Re(a b) written as in diabatic-Ehren.f:364
__device__ inline double re_mul(cplx a, cplx b) { return a.re * b.re - a.im * b.im; }

// Thread i of a workgroup owns row i for EH_JB consecutive columns j: its
// row-side bra and ket values (and IP, k_WH, V_shift) are loaded once, the
// column-side ones are the same for the whole workgroup.  A and K are written
// as coalesced column segments; rows and columns >= N are written as 0.
constexpr int EH_JB = 32;

__global__ __launch_bounds__(256) void ehrenfest_prep_kernel(EhrenfestArgs e) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= e.ld) return;
    const int j0 = blockIdx.y * EH_JB, j1 = min(e.ld, j0 + EH_JB);
    auto bra = [&](int r, int col) {
        cplx v = e.bra[(long)r * e.bra_stride + col];
        if (e.conj_bra) v.im = -v.im;
        return v;
    };
    auto ket = [&](int r, int col) { return e.ket[(long)r * e.ket_stride + col]; };
    const bool row = i < e.N, hole = e.bra_hl >= 0;
    cplx bie{0.0, 0.0}, bih{0.0, 0.0}, kie{0.0, 0.0}, kih{0.0, 0.0};
    double ipi = 0.0, kwi = 0.0, vsi = 0.0;
    if (row) {
        bie = bra(i, e.bra_el);
        kie = ket(i, e.ket_el);
        if (hole) {
            bih = bra(i, e.bra_hl);
            kih = ket(i, e.ket_hl);
        }
        if (!e.X) {
            ipi = e.ip[i];
            kwi = e.kwh[i];
            vsi = e.vshift[i];
        }
    }
    for (int j = j0; j < j1; ++j) {
        const long o = (long)j * e.ld + i;
        if (!row || j >= e.N) {
            e.A[o] = 0.0;
            e.K[o] = 0.0;
            continue;
        }
        double rij = re_mul(ket(j, e.ket_el), bie);
        double rji = re_mul(kie, bra(j, e.bra_el));
        if (hole) {
            rij -= re_mul(ket(j, e.ket_hl), bih);
            rji -= re_mul(kih, bra(j, e.bra_hl));
        }
        const double a = (rij + rji) * 0.5;
        double x;
        if (e.X) x = e.X[(long)j * e.ldx + i];
        else if (i == j) x = ipi + vsi;
        else x = huckel_x(ipi, e.ip[j], kwi, e.kwh[j], vsi, e.vshift[j]);
        e.A[o] = a;
        e.K[o] = x * a;
    }
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

int launch_ehrenfest(const EhrenfestArgs& e, const double* Hp, LibHandle* lib, hipStream_t st) {
    hipLaunchKernelGGL(ehrenfest_prep_kernel, dim3((e.ld + 255) / 256, (e.ld + EH_JB - 1) / EH_JB), dim3(256), 0, st,
                       e);
    if (lib) return lib_dgemm_nn(lib, e.N, e.N, e.N, -1.0, Hp, e.ld, e.A, e.ld, 1.0, e.K, e.ld);
    Gemm g{};
    g.M = e.N; g.N = e.N; g.K = e.N;
    g.A = Hp; g.lda = e.ld;
    g.B = e.A; g.ldb = e.ld;
    g.C = e.K; g.ldc = e.ld;
    g.alpha = -1.0; g.beta = 1.0; g.kmode = K_FULL; g.batch = 1;
    launch_gemm(g, st);
    return 0;
}

}  // namespace dh
