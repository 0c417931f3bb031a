// Dense FP64 algebra for the orthogonalisation S = L L^T, L^-1 and
// H_orth = L^-1 h L^-T, on gfx950 matrix cores (v_mfma_f64_16x16x4f64).
//
// Replaces the reference's S^-1 (dsytrf/dsytri on the CPU, MAGMA
// dgetrf/dgetri on the GPU: GPU_Interface.cpp:880-949) and H' = S^-1 h
// (xPU_dsymm / cublasDsymm: GPU_Interface.cpp:574-627, Taylor_gpu.cpp:695).
// All matrices are column-major with a leading dimension padded to a multiple
// of 64; the padding of S carries an identity block, so every blocked
// routine runs on whole 64 x 64 tiles.
#include "common.hpp"
#include "kernels.hpp"

namespace dh {

constexpr int GK = 16;  // k-depth of one LDS stage
constexpr int PAD = TILE + 1;

// C = alpha * opA(A) opB(B) + beta * C on 64 x 64 tiles; 256 threads = 4 waves
// in a 2 x 2 arrangement, each wave owning a 32 x 32 quadrant = 2 x 2 MFMA
// 16 x 16 blocks.  MFMA f64 16x16x4 operand layout: lane l holds A[l%16][l/16]
// and B[l/16][l%16]; accumulator value v of lane l is D[l/16 + 4*v][l%16]
// (f64 layout: the four values of a lane are strided by 4 rows).
__global__ __launch_bounds__(256) void gemm_f64_kernel(Gemm g) {
    __shared__ double As[GK][PAD];
    __shared__ double Bs[GK][PAD];
    const int bi = blockIdx.x, bj = blockIdx.y, z = blockIdx.z;
    if (g.lower_only && bi < bj) return;
    const int i0 = bi * TILE, j0 = bj * TILE;
    const double* A = g.A + z * g.strideA;
    const double* B = g.B + z * g.strideB;
    double* C = g.C + z * g.strideC;
    int klo = 0, khi = g.K;
    if (g.kmode == K_BELOW_ROWTILE) khi = min(g.K, i0 + TILE);
    else if (g.kmode == K_BELOW_COLTILE) khi = min(g.K, j0 + TILE);
    else if (g.kmode == K_FROM_ROWTILE) klo = i0;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wi = (w >> 1) * 32, wj = (w & 1) * 32;
    v4d acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = (v4d){0.0, 0.0, 0.0, 0.0};

    for (int k0 = klo; k0 < khi; k0 += GK) {
        // stage A tile: rows i0..i0+63, k0..k0+15
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int ii, kk;
            if (!g.transA) {
                ii = t & 63;
                kk = (t >> 6) + 4 * r;
            } else {
                kk = t & 15;
                ii = (t >> 4) + 16 * r;
            }
            const int gi = i0 + ii, gk = k0 + kk;
            double v = 0.0;
            if (gi < g.M && gk < khi) v = g.transA ? A[gk + g.lda * (long)gi] : A[gi + g.lda * (long)gk];
            As[kk][ii] = v;
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int jj, kk;
            if (!g.transB) {
                kk = t & 15;
                jj = (t >> 4) + 16 * r;
            } else {
                jj = t & 63;
                kk = (t >> 6) + 4 * r;
            }
            const int gj = j0 + jj, gk = k0 + kk;
            double v = 0.0;
            if (gj < g.N && gk < khi) v = g.transB ? B[gj + g.ldb * (long)gk] : B[gk + g.ldb * (long)gj];
            Bs[kk][jj] = v;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < GK; ks += 4) {
            double av[2], bv[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) av[a] = As[ks + (lane >> 4)][wi + a * 16 + (lane & 15)];
#pragma unroll
            for (int b = 0; b < 2; ++b) bv[b] = Bs[ks + (lane >> 4)][wj + b * 16 + (lane & 15)];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int gi = i0 + wi + a * 16 + (lane >> 4) + 4 * v;
                const int gj = j0 + wj + b * 16 + (lane & 15);
                if (gi < g.M && gj < g.N) {
                    const long idx = gi + g.ldc * (long)gj;
                    const double old = g.beta == 0.0 ? 0.0 : g.beta * C[idx];
                    C[idx] = g.alpha * acc[a][b][v] + old;
                }
            }
}

void launch_gemm(const Gemm& g, hipStream_t st) {
    if (g.M <= 0 || g.N <= 0) return;
    dim3 grid((g.M + TILE - 1) / TILE, (g.N + TILE - 1) / TILE, g.batch > 0 ? g.batch : 1);
    hipLaunchKernelGGL(gemm_f64_kernel, grid, dim3(256), 0, st, g);
}

// Factor the 64 x 64 diagonal block at (j0, j0) in place (lower) and write its
// inverse to Dinv[blk].  One wavefront; the block lives in LDS.
__global__ __launch_bounds__(64) void potrf_diag_kernel(double* A, int ld, int j0, double* Dinv, int* info) {
    __shared__ double Ls[TILE][PAD];
    __shared__ double Xs[TILE][PAD];
    const int t = threadIdx.x;
    for (int c = 0; c < TILE; ++c) Ls[t][c] = A[(long)(j0 + t) + (long)ld * (j0 + c)];
    __syncthreads();
    for (int j = 0; j < TILE; ++j) {
        double d = Ls[j][j];
        if (!(d > 0.0)) {
            if (t == 0 && *info == 0) *info = j0 + j + 1;
            d = 1.0;
        }
        const double djj = sqrt(d);
        __syncthreads();
        if (t == j) Ls[j][j] = djj;
        if (t > j) Ls[t][j] = Ls[t][j] / djj;
        __syncthreads();
        if (t > j) {
            const double lij = Ls[t][j];
            for (int k = j + 1; k <= t; ++k) Ls[t][k] -= lij * Ls[k][j];
        }
        __syncthreads();
    }
    // inverse of the lower-triangular block: thread t owns column t
    for (int i = 0; i < TILE; ++i) {
        double x = 0.0;
        if (i == t) x = 1.0 / Ls[i][i];
        else if (i > t) {
            double s = 0.0;
            for (int k = t; k < i; ++k) s += Ls[i][k] * Xs[k][t];
            x = -s / Ls[i][i];
        }
        Xs[i][t] = x;
        __syncthreads();
    }
    const int blk = j0 / TILE;
    for (int c = 0; c < TILE; ++c) {
        A[(long)(j0 + t) + (long)ld * (j0 + c)] = t >= c ? Ls[t][c] : 0.0;
        Dinv[(long)blk * TILE * TILE + t + (long)TILE * c] = Xs[t][c];
    }
}

// zero the strictly upper part of the panel columns j0..j0+63 above the diagonal block
__global__ void zero_upper_panel_kernel(double* A, int ld, int j0) {
    const int c = blockIdx.x;  // column within the panel
    for (int i = threadIdx.x; i < j0; i += blockDim.x) A[(long)i + (long)ld * (j0 + c)] = 0.0;
}

void launch_cholesky(double* A, int ld, double* Dinv, int* info, hipStream_t st) {
    const int nb = ld / TILE;
    for (int b = 0; b < nb; ++b) {
        const int j0 = b * TILE;
        hipLaunchKernelGGL(potrf_diag_kernel, dim3(1), dim3(64), 0, st, A, ld, j0, Dinv, info);
        if (j0 > 0) hipLaunchKernelGGL(zero_upper_panel_kernel, dim3(TILE), dim3(256), 0, st, A, ld, j0);
        const int rest = ld - j0 - TILE;
        if (rest <= 0) continue;
        // panel: A21 <- A21 * L11^-T (Dinv holds L11^-1).  In place is safe:
        // N = 64 gives one column tile per row tile, and a tile reads all 64
        // k-columns of its own rows before it writes them.
        Gemm g{};
        g.M = rest; g.N = TILE; g.K = TILE;
        g.A = A + (j0 + TILE) + (long)ld * j0; g.lda = ld;
        g.B = Dinv + (long)b * TILE * TILE; g.ldb = TILE; g.transB = 1;
        g.C = A + (j0 + TILE) + (long)ld * j0; g.ldc = ld;
        g.alpha = 1.0; g.beta = 0.0; g.batch = 1;
        launch_gemm(g, st);
        // trailing update (lower tiles only): A22 -= A21 A21^T
        Gemm u{};
        u.M = rest; u.N = rest; u.K = TILE;
        u.A = A + (j0 + TILE) + (long)ld * j0; u.lda = ld;
        u.B = A + (j0 + TILE) + (long)ld * j0; u.ldb = ld; u.transB = 1;
        u.C = A + (j0 + TILE) + (long)ld * (j0 + TILE); u.ldc = ld;
        u.alpha = -1.0; u.beta = 1.0; u.lower_only = 1; u.batch = 1;
        launch_gemm(u, st);
    }
}

// L^-1 by blocked forward substitution over block rows:
//   X[i, 0:i] = -Dinv_i * (L[i, 0:i] X[0:i, 0:i]),  X[i, i] = Dinv_i.
// work: 64 x ld scratch.
__global__ void place_diag_blocks_kernel(double* X, int ld, const double* Dinv) {
    const int b = blockIdx.x;
    for (int e = threadIdx.x; e < TILE * TILE; e += blockDim.x) {
        const int r = e % TILE, c = e / TILE;
        X[(long)(b * TILE + r) + (long)ld * (b * TILE + c)] = Dinv[(long)b * TILE * TILE + e];
    }
}
__global__ void fill_zero_kernel(double* X, long n) {
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) X[t] = 0.0;
}

void launch_trtri_lower(const double* L, double* X, const double* Dinv, int ld, double* work, hipStream_t st) {
    const long n = (long)ld * ld;
    hipLaunchKernelGGL(fill_zero_kernel, dim3(4096), dim3(256), 0, st, X, n);
    const int nb = ld / TILE;
    hipLaunchKernelGGL(place_diag_blocks_kernel, dim3(nb), dim3(256), 0, st, X, ld, Dinv);
    for (int b = 1; b < nb; ++b) {
        const int i0 = b * TILE;
        // Y = L[i0:i0+64, 0:i0] * X[0:i0, 0:i0]   (X lower: k >= j, i.e. rows of X from the column tile on)
        Gemm g{};
        g.M = TILE; g.N = i0; g.K = i0;
        g.A = L + i0; g.lda = ld;
        g.B = X; g.ldb = ld;
        g.C = work; g.ldc = TILE;
        g.alpha = 1.0; g.beta = 0.0; g.batch = 1;
        launch_gemm(g, st);
        // X[i0:i0+64, 0:i0] = -Dinv_b * Y
        Gemm h{};
        h.M = TILE; h.N = i0; h.K = TILE;
        h.A = Dinv + (long)b * TILE * TILE; h.lda = TILE;
        h.B = work; h.ldb = TILE;
        h.C = X + i0; h.ldc = ld;
        h.alpha = -1.0; h.beta = 0.0; h.batch = 1;
        launch_gemm(h, st);
    }
}

__global__ void symmetrize_from_lower_kernel(double* A, int ld) {
    // tile-transposing copy of the strictly lower part into the upper part
    __shared__ double tile[32][33];
    const int bx = blockIdx.x, by = blockIdx.y;
    if (bx > by) return;  // source tile (row block by, col block bx) in the lower part
    const int tx = threadIdx.x, ty = threadIdx.y;
    for (int r = ty; r < 32; r += blockDim.y) {
        const int i = by * 32 + r, j = bx * 32 + tx;
        tile[r][tx] = A[(long)i + (long)ld * j];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += blockDim.y) {
        // destination (row j, col i) with j = bx*32 + r, i = by*32 + tx
        const int j = bx * 32 + r, i = by * 32 + tx;
        if (i > j) A[(long)j + (long)ld * i] = tile[tx][r];
    }
}

void launch_symmetrize_from_lower(double* A, int ld, hipStream_t st) {
    const int nb = ld / 32;
    hipLaunchKernelGGL(symmetrize_from_lower_kernel, dim3(nb, nb), dim3(32, 8), 0, st, A, ld);
}

__global__ void copy_kernel(const double* s, double* d, long n) {
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) d[t] = s[t];
}
void launch_copy_matrix(const double* src, double* dst, long n, hipStream_t st) {
    hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, st, src, dst, n);
}

// S <- identity, h <- 0 over the whole ld x ld buffer (the host block is copied in afterwards).
__global__ void identity_pad_kernel(double* S, double* h, int ld) {
    const long total = (long)ld * ld;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const int i = (int)(t % ld), j = (int)(t / ld);
        S[t] = i == j ? 1.0 : 0.0;
        h[t] = 0.0;
    }
}
void launch_identity_pad(double* S, double* h, int N, int ld, hipStream_t st) {
    (void)N;
    hipLaunchKernelGGL(identity_pad_kernel, dim3(4096), dim3(256), 0, st, S, h, ld);
}

}  // namespace dh
