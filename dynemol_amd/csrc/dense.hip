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
// Tile of a workgroup.  Workgroups start in blockIdx order (x fastest); with a
// k range that grows with the row (column) tile, the rows (columns) are taken
// in reverse so that the heaviest tiles start first and the last ones to
// start are short.  Each tile's arithmetic is the same either way.
__device__ __forceinline__ void gemm_tile_of(const Gemm& g, int& bi, int& bj) {
    bi = blockIdx.x;
    bj = blockIdx.y;
    if (g.kmode == K_BELOW_ROWTILE) bi = gridDim.x - 1 - bi;
    else if (g.kmode == K_BELOW_COLTILE) bj = gridDim.y - 1 - bj;
}

__global__ __launch_bounds__(256) void gemm_f64_kernel(Gemm g) {
    __shared__ double As[GK][PAD];
    __shared__ double Bs[GK][PAD];
    int bi, bj;
    gemm_tile_of(g, bi, bj);
    const int z = blockIdx.z;
    if (g.lower_only && bi < bj) return;
    const int i0 = bi * TILE, j0 = bj * TILE;
    const double* A = g.A + z * g.strideA;
    const double* B = g.B + z * g.strideB;
    double* C = g.C + z * g.strideC;
    int klo = 0, khi = g.K;
    if (g.kmode == K_BELOW_ROWTILE) khi = min(g.K, i0 + TILE);
    else if (g.kmode == K_BELOW_COLTILE) khi = min(g.K, j0 + TILE);
    else if (g.kmode == K_FROM_ROWTILE) klo = i0;
    else if (g.kmode == K_FROM_COLTILE) klo = j0;
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

// The same product on 128 x 128 tiles for large M and N: 4 waves in a 2 x 2
// arrangement, each wave a 64 x 64 quadrant = 4 x 4 MFMA blocks; a 16-deep k
// stage is double-buffered in LDS and the next stage is fetched into registers
// while the current one feeds the matrix cores (one barrier per stage).  The
// MFMA is issued with the operands swapped (D^T = B^T A^T), so a lane's four
// results lie in four columns and each column's 16 rows are contiguous in C.
// lower_only: tiles with row-tile < col-tile are skipped at the 64 x 64 level,
// as in the 64 x 64 kernel.
constexpr int G2_T = 128, G2_BK = 16, G2_PAD = G2_T + 4;

// One 128 x 128 output tile (bi, bj) of batch z; the LDS stages are the caller's.
template <int TA, int TB>
__device__ __forceinline__ void gemm128_tile(const Gemm& g, int bi, int bj, int z, double (*As)[G2_BK][G2_PAD],
                                             double (*Bs)[G2_BK][G2_PAD]) {
    if (g.lower_only && bi < bj) return;
    const int i0 = bi * G2_T, j0 = bj * G2_T;
    const double* A = g.A + z * g.strideA;
    const double* B = g.B + z * g.strideB;
    double* C = g.C + z * g.strideC;
    int klo = 0, khi = g.K;
    if (g.kmode == K_BELOW_ROWTILE) khi = min(g.K, i0 + G2_T);
    else if (g.kmode == K_BELOW_COLTILE) khi = min(g.K, j0 + G2_T);
    else if (g.kmode == K_FROM_ROWTILE) klo = i0;
    else if (g.kmode == K_FROM_COLTILE) klo = j0;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wi = (w >> 1) * 64, wj = (w & 1) * 64;
    v4d acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = (v4d){0.0, 0.0, 0.0, 0.0};
    // C <- C - A B (alpha = -1, beta = 1): C is read into the accumulators
    // before the k loop, where its latency overlaps the operand loads, and the
    // products are subtracted (A negated on its way out of LDS, exactly)
    const bool cacc = g.c_acc != 0;
    if (cacc) {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int gi = i0 + wi + a * 16 + (lane & 15);
                    const int gj = j0 + wj + b * 16 + (lane >> 4) + 4 * v;
                    acc[a][b][v] = (gi < g.M && gj < g.N) ? C[gi + g.ldc * (long)gj] : 0.0;
                }
    }

    double ra[8], rb[8];
    auto gload = [&](int k0) {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int ii = TA ? (t >> 4) + 16 * r : (t & 127);
            const int kk = TA ? (t & 15) : (t >> 7) + 2 * r;
            const int gi = i0 + ii, gk = k0 + kk;
            double v = 0.0;
            if (gi < g.M && gk < khi) v = TA ? A[gk + g.lda * (long)gi] : A[gi + g.lda * (long)gk];
            ra[r] = v;
        }
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int jj = TB ? (t & 127) : (t >> 4) + 16 * r;
            const int kk = TB ? (t >> 7) + 2 * r : (t & 15);
            const int gj = j0 + jj, gk = k0 + kk;
            double v = 0.0;
            if (gj < g.N && gk < khi) v = TB ? B[gj + g.ldb * (long)gk] : B[gk + g.ldb * (long)gj];
            rb[r] = v;
        }
    };
    auto swrite = [&](int s) {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int ii = TA ? (t >> 4) + 16 * r : (t & 127);
            const int kk = TA ? (t & 15) : (t >> 7) + 2 * r;
            As[s][kk][ii] = ra[r];
        }
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int jj = TB ? (t & 127) : (t >> 4) + 16 * r;
            const int kk = TB ? (t >> 7) + 2 * r : (t & 15);
            Bs[s][kk][jj] = rb[r];
        }
    };
    int s = 0;
    if (klo < khi) {
        gload(klo);
        swrite(0);
    }
    __syncthreads();
    for (int k0 = klo; k0 < khi; k0 += G2_BK) {
        const bool more = k0 + G2_BK < khi;
        if (more) gload(k0 + G2_BK);
#pragma unroll
        for (int ks = 0; ks < G2_BK; ks += 4) {
            double av[4], bv[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const double x = As[s][ks + (lane >> 4)][wi + a * 16 + (lane & 15)];
                av[a] = cacc ? -x : x;
            }
#pragma unroll
            for (int b = 0; b < 4; ++b) bv[b] = Bs[s][ks + (lane >> 4)][wj + b * 16 + (lane & 15)];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[b], av[a], acc[a][b], 0, 0, 0);
        }
        if (more) swrite(s ^ 1);
        __syncthreads();
        s ^= 1;
    }
    if (g.lower_only && (i0 + wi) / TILE < (j0 + wj) / TILE) return;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int gi = i0 + wi + a * 16 + (lane & 15);
                const int gj = j0 + wj + b * 16 + (lane >> 4) + 4 * v;
                if (gi < g.M && gj < g.N) {
                    const long idx = gi + g.ldc * (long)gj;
                    if (cacc) {
                        C[idx] = acc[a][b][v];
                    } else {
                        const double old = g.beta == 0.0 ? 0.0 : g.beta * C[idx];
                        C[idx] = g.alpha * acc[a][b][v] + old;
                    }
                }
            }
}

// A workgroup computes one tile, or with g.pair two tiles whose k ranges add up
// to the same length (row tiles i and n-1-i of a triangular k range, or the
// same for columns): every workgroup then does equal work, which a one-tile
// grid cannot give when the heavy tiles land on a few XCDs.
template <int TA, int TB>
__global__ __launch_bounds__(256, 2) void gemm128_f64_kernel(Gemm g) {
    __shared__ double As[2][G2_BK][G2_PAD];
    __shared__ double Bs[2][G2_BK][G2_PAD];
    const int z = blockIdx.z;
    int bi, bj;
    gemm_tile_of(g, bi, bj);
    int npass = 1;
    if (g.pair) {
        const int rows = g.pair == 1;
        const int n = rows ? (g.M + G2_T - 1) / G2_T : (g.N + G2_T - 1) / G2_T;
        bi = blockIdx.x;
        bj = blockIdx.y;
        npass = (rows ? n - 1 - bi : n - 1 - bj) != (rows ? bi : bj) ? 2 : 1;
    }
#pragma unroll 1
    for (int pass = 0; pass < npass; ++pass) {
        if (pass == 1) {
            __syncthreads();  // the first tile's LDS readers are done
            if (g.pair == 1) bi = (g.M + G2_T - 1) / G2_T - 1 - bi;
            else bj = (g.N + G2_T - 1) / G2_T - 1 - bj;
        }
        gemm128_tile<TA, TB>(g, bi, bj, z, As, Bs);
    }
}

void launch_gemm(const Gemm& g, hipStream_t st) {
    if (g.M <= 0 || g.N <= 0) return;
    const unsigned nz = g.batch > 0 ? g.batch : 1;
    if (g.M >= 2 * G2_T && g.N >= 2 * G2_T) {
        Gemm p = g;
        dim3 grid((g.M + G2_T - 1) / G2_T, (g.N + G2_T - 1) / G2_T, nz);
        p.pair = 0;
        p.c_acc = g.alpha == -1.0 && g.beta == 1.0;
        if (!g.lower_only && (g.kmode == K_BELOW_ROWTILE || g.kmode == K_FROM_ROWTILE)) {
            p.pair = 1;
            grid.x = (grid.x + 1) / 2;
        } else if (!g.lower_only && (g.kmode == K_BELOW_COLTILE || g.kmode == K_FROM_COLTILE)) {
            p.pair = 2;
            grid.y = (grid.y + 1) / 2;
        }
        if (!g.transA && !g.transB) hipLaunchKernelGGL((gemm128_f64_kernel<0, 0>), grid, dim3(256), 0, st, p);
        else if (!g.transA) hipLaunchKernelGGL((gemm128_f64_kernel<0, 1>), grid, dim3(256), 0, st, p);
        else if (!g.transB) hipLaunchKernelGGL((gemm128_f64_kernel<1, 0>), grid, dim3(256), 0, st, p);
        else hipLaunchKernelGGL((gemm128_f64_kernel<1, 1>), grid, dim3(256), 0, st, p);
        return;
    }
    dim3 grid((g.M + TILE - 1) / TILE, (g.N + TILE - 1) / TILE, nz);
    hipLaunchKernelGGL(gemm_f64_kernel, grid, dim3(256), 0, st, g);
}

// Factor the 64 x 64 diagonal block at (j0, j0) in place (lower) and write its
// inverse to Dinv[blk].  1024 threads, the block in LDS (column-major, padded),
// blocked by 16 so that only 20 barriers separate the serial parts:
//   factor, for each 16-column sub-panel k0:
//     1. wave 0 factors the 16 x 16 diagonal sub-block, lane r holding row r in
//        registers and taking column entries of other rows by shuffles;
//     2. one thread per row below solves L_i,panel = A_i,panel L_D^-T by forward
//        substitution over the 16 columns;
//     3. all threads apply the rank-16 update to the lower part below;
//   inverse X = L^-1:
//     4. waves 0..3 invert the four 16 x 16 diagonal sub-blocks (lane c solves
//        column c by forward substitution);
//     5. sub-diagonal d = 1, 2, 3: T_IJ = sum_{K=J..I-1} L_IK X_KJ, then
//        X_IJ = -X_II T_IJ, for all sub-blocks I - J = d at once.
// The first non-positive pivot is reported as in LAPACK's potrf (1-based).
constexpr int PD_P = TILE + 1;  // padded LDS column
constexpr int PD_B = 16;        // sub-block size
__global__ __launch_bounds__(1024) void potrf_diag_kernel(double* A, int ld, int j0, double* Dinv, int* info) {
    __shared__ double Ls[TILE * PD_P];  // Ls[c * PD_P + r] = block (r, c)
    __shared__ double Xs[TILE * PD_P];  // inverse, same layout
    __shared__ double Ts[TILE * PD_P];  // products of step 5
    __shared__ int bad;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t == 0) bad = TILE + 1;
    for (int e = t; e < TILE * TILE; e += 1024) {
        const int r = e % TILE, c = e / TILE;
        Ls[c * PD_P + r] = A[(long)(j0 + r) + (long)ld * (j0 + c)];
        Xs[c * PD_P + r] = 0.0;
    }
    __syncthreads();
    for (int k0 = 0; k0 < TILE; k0 += PD_B) {
        if (w == 0) {  // 1. diagonal sub-block (wave-uniform branch: no barrier inside)
            const int r = lane < PD_B ? lane : 0;
            double a[PD_B];
#pragma unroll
            for (int c = 0; c < PD_B; ++c) a[c] = Ls[(k0 + c) * PD_P + k0 + r];
#pragma unroll
            for (int j = 0; j < PD_B; ++j) {
                const double d = __shfl(a[j], j);
                const bool ok = d > 0.0;
                if (!ok && lane == 0) atomicMin(&bad, k0 + j + 1);
                const double sd = ok ? sqrt(d) : 1.0;
                a[j] = lane == j ? sd : (lane > j ? a[j] / sd : a[j]);
#pragma unroll
                for (int c = j + 1; c < PD_B; ++c) {
                    const double lcj = __shfl(a[j], c);  // L_cj from row c
                    if (lane >= c) a[c] -= a[j] * lcj;
                }
            }
            if (lane < PD_B) {
#pragma unroll
                for (int c = 0; c < PD_B; ++c) Ls[(k0 + c) * PD_P + k0 + lane] = c <= lane ? a[c] : 0.0;
            }
        }
        __syncthreads();
        const int n2 = TILE - k0 - PD_B;  // rows below the sub-block
        if (n2 <= 0) break;
        if (t < n2) {  // 2. sub-panel solve, one row per thread
            const int i = k0 + PD_B + t;
            double x[PD_B];
#pragma unroll
            for (int c = 0; c < PD_B; ++c) {
                double v = Ls[(k0 + c) * PD_P + i];
#pragma unroll
                for (int k = 0; k < c; ++k) v -= x[k] * Ls[(k0 + k) * PD_P + k0 + c];
                x[c] = v / Ls[(k0 + c) * PD_P + k0 + c];
            }
#pragma unroll
            for (int c = 0; c < PD_B; ++c) Ls[(k0 + c) * PD_P + i] = x[c];
        }
        __syncthreads();
        for (int e = t; e < n2 * n2; e += 1024) {  // 3. rank-16 update, lower part
            const int i = k0 + PD_B + e % n2, j = k0 + PD_B + e / n2;
            if (i < j) continue;
            double v = Ls[j * PD_P + i];
#pragma unroll
            for (int k = 0; k < PD_B; ++k) v -= Ls[(k0 + k) * PD_P + i] * Ls[(k0 + k) * PD_P + j];
            Ls[j * PD_P + i] = v;
        }
        __syncthreads();
    }
    if (w < TILE / PD_B && lane < PD_B) {  // 4. diagonal sub-block inverses
        const int b0 = PD_B * w, c = lane;
        double x[PD_B];
#pragma unroll
        for (int i = 0; i < PD_B; ++i) {
            double v = i == c ? 1.0 : 0.0;
#pragma unroll
            for (int k = 0; k < i; ++k) v -= Ls[(b0 + k) * PD_P + b0 + i] * x[k];
            x[i] = v / Ls[(b0 + i) * PD_P + b0 + i];
        }
#pragma unroll
        for (int i = 0; i < PD_B; ++i) Xs[(b0 + c) * PD_P + b0 + i] = x[i];
    }
    __syncthreads();
    constexpr int NBK = TILE / PD_B;
    for (int d = 1; d < NBK; ++d) {  // 5. off-diagonal sub-blocks by sub-diagonal
        const int nblk = NBK - d;
        for (int e = t; e < nblk * PD_B * PD_B; e += 1024) {
            const int J = e / (PD_B * PD_B), I = J + d, i = e % PD_B, j = (e / PD_B) % PD_B;
            double v = 0.0;
            for (int k = PD_B * J; k < PD_B * I; ++k) v += Ls[k * PD_P + PD_B * I + i] * Xs[(PD_B * J + j) * PD_P + k];
            Ts[(PD_B * J + j) * PD_P + PD_B * I + i] = v;
        }
        __syncthreads();
        for (int e = t; e < nblk * PD_B * PD_B; e += 1024) {
            const int J = e / (PD_B * PD_B), I = J + d, i = e % PD_B, j = (e / PD_B) % PD_B;
            double v = 0.0;
#pragma unroll
            for (int k = 0; k < PD_B; ++k)
                v -= Xs[(PD_B * I + k) * PD_P + PD_B * I + i] * Ts[(PD_B * J + j) * PD_P + PD_B * I + k];
            Xs[(PD_B * J + j) * PD_P + PD_B * I + i] = v;
        }
        __syncthreads();
    }
    const int blk = j0 / TILE;
    for (int e = t; e < TILE * TILE; e += 1024) {
        const int r = e % TILE, c = e / TILE;
        A[(long)(j0 + r) + (long)ld * (j0 + c)] = r >= c ? Ls[c * PD_P + r] : 0.0;
        Dinv[(long)blk * TILE * TILE + r + (long)TILE * c] = r >= c ? Xs[c * PD_P + r] : 0.0;
    }
    if (t == 0 && bad <= TILE && *info == 0) *info = j0 + bad;
}

// zero the blocks above the diagonal blocks (column j: rows 0 .. 64 floor(j / 64) - 1),
// once the factorization is done: the trailing updates run on lower tiles of
// 128 and leave values in some of these blocks, and nothing reads them before
__global__ void zero_upper_blocks_kernel(double* A, int ld) {
    const int j = blockIdx.x;
    const int top = (j / TILE) * TILE;
    for (int i = threadIdx.x; i < top; i += blockDim.x) A[(long)i + (long)ld * j] = 0.0;
}

// Right-looking blocked Cholesky in panels of CHOL_PANEL 64-column blocks:
// within a panel, block b is factored (potrf_diag_kernel), its column below is
// solved against it, and the rank-64 update reaches only the panel's later
// columns; the matrix right of the panel then takes one rank-256 update.  The
// trailing matrix is read and written once per 256 columns instead of once per
// 64 (at N = 4096 about 1.4 GB of update traffic instead of 5.7 GB), and its
// GEMMs run with K = 256.
constexpr int CHOL_PANEL = 4;

void launch_cholesky(double* A, int ld, double* Dinv, int* info, hipStream_t st) {
    const int nb = ld / TILE;
    for (int p0 = 0; p0 < nb; p0 += CHOL_PANEL) {
        const int pe = std::min(nb, p0 + CHOL_PANEL);  // blocks p0 .. pe-1
        for (int b = p0; b < pe; ++b) {
            const int j0 = b * TILE;
            hipLaunchKernelGGL(potrf_diag_kernel, dim3(1), dim3(1024), 0, st, A, ld, j0, Dinv, info);
            const int rest = ld - j0 - TILE;
            if (rest <= 0) continue;
            // column: A21 <- A21 * L11^-T (Dinv holds L11^-1).  In place is
            // safe: N = 64 gives one column tile per row tile, and a tile reads
            // all 64 k-columns of its own rows before it writes them.
            Gemm g{};
            g.M = rest; g.N = TILE; g.K = TILE;
            g.A = A + (j0 + TILE) + (long)ld * j0; g.lda = ld;
            g.B = Dinv + (long)b * TILE * TILE; g.ldb = TILE; g.transB = 1;
            g.C = A + (j0 + TILE) + (long)ld * j0; g.ldc = ld;
            g.alpha = 1.0; g.beta = 0.0; g.batch = 1;
            launch_gemm(g, st);
            // the panel's later columns (lower tiles only): A22[:, :w] -= A21 A21[:w]^T
            const int w = pe * TILE - (j0 + TILE);
            if (w <= 0) continue;
            Gemm u{};
            u.M = rest; u.N = w; u.K = TILE;
            u.A = A + (j0 + TILE) + (long)ld * j0; u.lda = ld;
            u.B = A + (j0 + TILE) + (long)ld * j0; u.ldb = ld; u.transB = 1;
            u.C = A + (j0 + TILE) + (long)ld * (j0 + TILE); u.ldc = ld;
            u.alpha = -1.0; u.beta = 1.0; u.lower_only = 1; u.batch = 1;
            launch_gemm(u, st);
        }
        // right of the panel (lower tiles only): A33 -= L31 L31^T, K = the panel's width
        const int t0 = pe * TILE, rest = ld - t0;
        if (rest <= 0) continue;
        Gemm u{};
        u.M = rest; u.N = rest; u.K = (pe - p0) * TILE;
        u.A = A + t0 + (long)ld * (p0 * TILE); u.lda = ld;
        u.B = A + t0 + (long)ld * (p0 * TILE); u.ldb = ld; u.transB = 1;
        u.C = A + t0 + (long)ld * t0; u.ldc = ld;
        u.alpha = -1.0; u.beta = 1.0; u.lower_only = 1; u.batch = 1;
        launch_gemm(u, st);
    }
    if (nb > 1) hipLaunchKernelGGL(zero_upper_blocks_kernel, dim3(ld), dim3(256), 0, st, A, ld);
}

// L^-1 by recursive doubling over the 64 x 64 diagonal blocks (whose inverses
// Dinv the Cholesky left behind).  A node of the block hierarchy splits as
//   L = [L11 0; L21 L22]  ->  L^-1 = [X11 0; -X22 L21 X11  X22],
// and every node of one level is independent, so a level is two batched
// GEMMs (T = L21 X11 with X11 lower: k >= column; X21 = -X22 T with X22
// lower: k <= row).  A level whose last node has a short right child runs
// that node with its own sizes.  work: ld * ld / 2 doubles.
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

static void trtri_level(const double* L, double* X, int ld, double* work, int r0, int n1, int n2, int count,
                        long stride, hipStream_t st) {
    // T (n2 x n1) = L21 X11
    Gemm g{};
    g.M = n2; g.N = n1; g.K = n1;
    g.A = L + (r0 + n1) + (long)ld * r0; g.lda = ld; g.strideA = stride;
    g.B = X + r0 + (long)ld * r0; g.ldb = ld; g.strideB = stride;
    g.C = work; g.ldc = n2; g.strideC = (long)n2 * n1;
    g.alpha = 1.0; g.beta = 0.0; g.kmode = K_FROM_COLTILE; g.batch = count;
    launch_gemm(g, st);
    // X21 = -X22 T
    Gemm h{};
    h.M = n2; h.N = n1; h.K = n2;
    h.A = X + (r0 + n1) + (long)ld * (r0 + n1); h.lda = ld; h.strideA = stride;
    h.B = work; h.ldb = n2; h.strideB = (long)n2 * n1;
    h.C = X + (r0 + n1) + (long)ld * r0; h.ldc = ld; h.strideC = stride;
    h.alpha = -1.0; h.beta = 0.0; h.kmode = K_BELOW_ROWTILE; h.batch = count;
    launch_gemm(h, st);
}

void launch_trtri_lower(const double* L, double* X, const double* Dinv, int ld, double* work, hipStream_t st) {
    const long n = (long)ld * ld;
    hipLaunchKernelGGL(fill_zero_kernel, dim3(4096), dim3(256), 0, st, X, n);
    const int nb = ld / TILE;
    hipLaunchKernelGGL(place_diag_blocks_kernel, dim3(nb), dim3(256), 0, st, X, ld, Dinv);
    for (int hb = 1; hb < nb; hb *= 2) {  // child size in blocks
        const int b = hb * TILE;
        const int full = nb / (2 * hb);
        const long stride = 2L * b * (ld + 1);
        if (full > 0) trtri_level(L, X, ld, work, 0, b, b, full, stride, st);
        const int rem = nb - full * 2 * hb;
        if (rem > hb) trtri_level(L, X, ld, work, full * 2 * b, b, (rem - hb) * TILE, 1, 0, st);
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
