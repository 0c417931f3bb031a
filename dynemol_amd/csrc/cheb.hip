// Chebyshev exp(-i H dt) propagation kernels (FP64), spectral bounds and
// observables.
//
// The reference propagates with an adaptive-tau Taylor series, streaming H'
// twice per term (bra and ket GEMVs: Taylor.f:90-100, Taylor_gpu.cpp:411-416)
// with a host synchronisation per term.  Here every wavepacket's bra and ket
// are right-hand-side columns of one real-symmetric H_orth, and one kernel
// launch per Chebyshev term streams H_orth once for all of them:
//
//   T_0 = phi, T_1 = Hh phi, T_k = 2 Hh T_{k-1} - T_{k-2},  Hh = (H - c)/r
//   phi(t + dt) = sum_k a_k T_k,  a_k = (2 - delta_k0) (-i)^k J_k(r dt) e^{-i c dt}
//
// matvec_kernel: y_i = sum_j M[j + ld*i] x_j (column i of the column-major
// matrix is contiguous, so lanes stream it coalesced).  Grid = column blocks x
// row splits; each workgroup leaves a partial sum per column, and the last
// workgroup of a column block (ticket counter) adds the partials in split
// order and applies the fused epilogue.  The summation order is fixed, so
// results are bitwise reproducible run to run.
#include "common.hpp"
#include "kernels.hpp"

#include <cmath>
#include <type_traits>
#include <utility>
#include <mutex>
#include <vector>

namespace dh {

// The accumulator of a one-step series for output element oi, given T_k and
// the term's input x: EPI_CHEB: acc += a_k T_k; EPI_CHEB_FIRST:
// acc = a_0 x + a_1 T_1.  A null a.acc skips it (the recurrences of a
// multi-step series, whose accumulators take the terms in panels).
__device__ __forceinline__ void acc_update(const MatvecArgs& a, long oi, cplx xv, cplx T) {
    if (!a.acc) return;
    if (a.epi == EPI_CHEB) {
        const cplx add = cmul(a.coef[a.k], T);
        cplx acc = a.acc[oi];
        acc.re += add.re;
        acc.im += add.im;
        a.acc[oi] = acc;
    } else {
        const cplx t0 = cmul(a.coef[0], xv), t1 = cmul(a.coef[1], T);
        a.acc[oi] = cplx{t0.re + t1.re, t0.im + t1.im};
    }
}

constexpr int MV_WAVES = 4;

template <int R, int CPW>
__global__ __launch_bounds__(256) void matvec_kernel(MatvecArgs a) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ncol = MV_WAVES * CPW;
    const int i_base = blockIdx.x * ncol + wave * CPW;
    const int s = blockIdx.y;
    const int rows_per = (((a.N + a.split - 1) / a.split) + 63) & ~63;
    int j_begin = s * rows_per;
    int j_end = min(a.N, j_begin + rows_per);
    if (a.tri == TRI_LOWER) j_begin = max(j_begin, (i_base / 64) * 64);  // j >= i (masked per element below)
    if (a.tri == TRI_UPPER) j_end = min(j_end, i_base + CPW);             // j <= i

    double yr[CPW][R], yi[CPW][R];
#pragma unroll
    for (int c = 0; c < CPW; ++c)
#pragma unroll
        for (int r = 0; r < R; ++r) {
            yr[c][r] = 0.0;
            yi[c][r] = 0.0;
        }
    const long ld = a.ld;
    int ic[CPW];
#pragma unroll
    for (int c = 0; c < CPW; ++c) ic[c] = min(i_base + c, a.N - 1);  // clamped: loads stay in bounds

    for (int j0 = j_begin; j0 < j_end; j0 += 128) {
        double m[2][CPW];
        double xr[2][R], xi[2][R];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int j = j0 + u * 64 + lane;
            const bool ok = j < j_end;
            const int jj = ok ? j : j_begin;
#pragma unroll
            for (int c = 0; c < CPW; ++c) {
                double v = a.M[jj + ld * ic[c]];
                const int i = i_base + c;
                bool keep = ok && i < a.N;
                if (a.tri == TRI_LOWER) keep = keep && (jj >= i);
                if (a.tri == TRI_UPPER) keep = keep && (jj <= i);
                m[u][c] = keep ? v : 0.0;
            }
            const cplx* xp = a.x + (long)jj * a.x_stride + a.x_off;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const cplx xv = xp[r];
                xr[u][r] = ok ? xv.re : 0.0;
                xi[u][r] = ok ? xv.im : 0.0;
            }
        }
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int c = 0; c < CPW; ++c)
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    yr[c][r] = fma(m[u][c], xr[u][r], yr[c][r]);
                    yi[c][r] = fma(m[u][c], xi[u][r], yi[c][r]);
                }
    }
#pragma unroll
    for (int c = 0; c < CPW; ++c)
#pragma unroll
        for (int r = 0; r < R; ++r) {
            yr[c][r] = wave_sum(yr[c][r]);
            yi[c][r] = wave_sum(yi[c][r]);
        }

    __shared__ int s_last;
    __shared__ cplx s_y[MV_WAVES * CPW * R];
    if (a.split > 1) {
        if (lane == 0) {
#pragma unroll
            for (int c = 0; c < CPW; ++c) {
                const int i = i_base + c;
                if (i < a.N)
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        a.partial[((long)s * a.N + i) * R + r] = cplx{yr[c][r], yi[c][r]};
            }
        }
        __threadfence();
        __syncthreads();
        if (threadIdx.x == 0) {
            const unsigned ticket = atomicAdd(&a.counters[blockIdx.x], 1u);
            s_last = (ticket == (unsigned)(a.split - 1));
        }
        __syncthreads();
        if (!s_last) return;
        __threadfence();
        // sum partials in split order
        for (int e = threadIdx.x; e < ncol * R; e += blockDim.x) {
            const int c = e / R, r = e % R;
            const int i = blockIdx.x * ncol + c;
            cplx y{0.0, 0.0};
            if (i < a.N)
                for (int ss = 0; ss < a.split; ++ss) {
                    const cplx p = a.partial[((long)ss * a.N + i) * R + r];
                    y.re += p.re;
                    y.im += p.im;
                }
            s_y[e] = y;
        }
        if (threadIdx.x == 0) a.counters[blockIdx.x] = 0u;
    } else {
        if (lane == 0)
#pragma unroll
            for (int c = 0; c < CPW; ++c)
#pragma unroll
                for (int r = 0; r < R; ++r) s_y[(wave * CPW + c) * R + r] = cplx{yr[c][r], yi[c][r]};
    }
    __syncthreads();
    for (int e = threadIdx.x; e < ncol * R; e += blockDim.x) {
        const int c = e / R, r = e % R;
        const int i = blockIdx.x * ncol + c;
        if (i >= a.N) continue;
        const cplx y = s_y[e];
        const long oi = (long)i * a.out_stride + a.out_off + r;
        if (a.epi == EPI_PLAIN) {
            a.out[oi] = y;
        } else {
            const cplx xv = a.x[(long)i * a.x_stride + a.x_off + r];
            cplx T{a.alpha * y.re + a.beta * xv.re, a.alpha * y.im + a.beta * xv.im};
            if (a.epi == EPI_CHEB) {
                const cplx x2 = a.x2[(long)i * a.x_stride + a.x_off + r];
                T.re -= x2.re;
                T.im -= x2.im;
            }
            acc_update(a, oi, xv, T);
            a.out[oi] = T;
        }
    }
}

// Column-block matvec: a workgroup of NW waves owns NW*CPW consecutive output
// columns over the whole j range, so no other workgroup contributes to them
// (no partial sums, no fences).  x (T_{k-1}) is staged through a
// double-buffered LDS tile of JC rows, stored as separate re/im planes per
// column so lanes read consecutive doubles; the next tile is fetched into
// registers while the current one is consumed.
// PIPE = 0: let the compiler schedule the tile; 1: read x one row ahead and
// fence the scheduler per row (bounds register use); 2: fence per row, no lookahead.
template <int R, int CPW, int NW, int JC, int PIPE = 0>
__device__ __forceinline__ void cols_body(const MatvecArgs& a, const int block) {
    static_assert(JC % 64 == 0 && JC >= 64, "tile height must be a multiple of 64");
    static_assert(CPW * R <= 64, "one lane per output in the epilogue");
    constexpr int NT = NW * 64;
    constexpr int TE = JC * R;                 // complex elements per tile
    constexpr int PER = (TE + NT - 1) / NT;    // complex elements per thread per tile
    constexpr int TU = JC / 64;                // rows per lane per tile
    __shared__ double xs[2][2 * R][JC];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int NC = NW * CPW;
    const int c0 = block * NC;
    const int i_base = c0 + wave * CPW;
    int jb = 0, je = a.N;
    if (a.tri == TRI_LOWER) jb = (c0 / JC) * JC;
    if (a.tri == TRI_UPPER) je = min(a.N, c0 + NC);
    const long ld = a.ld;
    int ic[CPW];
#pragma unroll
    for (int c = 0; c < CPW; ++c) ic[c] = min(i_base + c, a.N - 1);
    double yr[CPW][R], yi[CPW][R];
#pragma unroll
    for (int c = 0; c < CPW; ++c)
#pragma unroll
        for (int r = 0; r < R; ++r) {
            yr[c][r] = 0.0;
            yi[c][r] = 0.0;
        }
    cplx pre[PER];
    auto fetch = [&](int t0) {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = tid + q * NT;
            const int j = t0 + e / R, r = e % R;
            cplx v{0.0, 0.0};
            if (e < TE && j < je) v = a.x[(long)j * a.x_stride + a.x_off + r];
            pre[q] = v;
        }
    };
    auto stash = [&](int buf) {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = tid + q * NT;
            if (e < TE) {
                const int jl = e / R, r = e % R;
                xs[buf][2 * r][jl] = pre[q].re;
                xs[buf][2 * r + 1][jl] = pre[q].im;
            }
        }
    };
    if (jb < je) {
        fetch(jb);
        stash(0);
    }
    __syncthreads();
    int buf = 0;
    for (int t0 = jb; t0 < je; t0 += JC) {
        const bool more = t0 + JC < je;
        if (more) fetch(t0 + JC);
        double hv[TU][CPW];
#pragma unroll
        for (int u = 0; u < TU; ++u) {
            const int j = t0 + u * 64 + lane;
            const int jj = j < je ? j : jb;
#pragma unroll
            for (int c = 0; c < CPW; ++c) {
                const double v = a.M[jj + ld * ic[c]];
                const int i = i_base + c;
                bool keep = j < je && i < a.N;
                if (a.tri == TRI_LOWER) keep = keep && (j >= i);
                if (a.tri == TRI_UPPER) keep = keep && (j <= i);
                hv[u][c] = keep ? v : 0.0;
            }
        }
if constexpr (PIPE == 0) {
#pragma unroll
            for (int u = 0; u < TU; ++u) {
                const int jl = u * 64 + lane;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const double xr = xs[buf][2 * r][jl], xi = xs[buf][2 * r + 1][jl];
#pragma unroll
                    for (int c = 0; c < CPW; ++c) {
                        yr[c][r] = fma(hv[u][c], xr, yr[c][r]);
                        yi[c][r] = fma(hv[u][c], xi, yi[c][r]);
                    }
                }
            }
        } else {
            double xc[2 * R], xn[2 * R];
#pragma unroll
            for (int q = 0; q < 2 * R; ++q) xc[q] = xs[buf][q][lane];
#pragma unroll
            for (int u = 0; u < TU; ++u) {
                if constexpr (PIPE == 1) {
                    if (u + 1 < TU) {
#pragma unroll
                        for (int q = 0; q < 2 * R; ++q) xn[q] = xs[buf][q][(u + 1) * 64 + lane];
                    }
                } else if (u > 0) {
#pragma unroll
                    for (int q = 0; q < 2 * R; ++q) xc[q] = xs[buf][q][u * 64 + lane];
                }
#pragma unroll
                for (int r = 0; r < R; ++r)
#pragma unroll
                    for (int c = 0; c < CPW; ++c) {
                        yr[c][r] = fma(hv[u][c], xc[2 * r], yr[c][r]);
                        yi[c][r] = fma(hv[u][c], xc[2 * r + 1], yi[c][r]);
                    }
                __builtin_amdgcn_sched_barrier(0);
                if constexpr (PIPE == 1) {
#pragma unroll
                    for (int q = 0; q < 2 * R; ++q) xc[q] = xn[q];
                }
            }
        }
        if (more) stash(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
    // reduce over lanes; lane (c*R + r) then owns output (i_base + c, r)
    double my_re = 0.0, my_im = 0.0;
#pragma unroll
    for (int c = 0; c < CPW; ++c)
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const double sr = wave_sum(yr[c][r]);
            const double si = wave_sum(yi[c][r]);
            if (lane == c * R + r) {
                my_re = sr;
                my_im = si;
            }
        }
    if (lane >= CPW * R) return;
    const int c = lane / R, r = lane % R;
    const int i = i_base + c;
    if (i >= a.N) return;
    const cplx y{my_re, my_im};
    const long oi = (long)i * a.out_stride + a.out_off + r;
    if (a.epi == EPI_PLAIN) {
        a.out[oi] = y;
        return;
    }
    const cplx xv = a.x[(long)i * a.x_stride + a.x_off + r];
    cplx T{a.alpha * y.re + a.beta * xv.re, a.alpha * y.im + a.beta * xv.im};
    if (a.epi == EPI_CHEB) {
        const cplx x2 = a.x2[(long)i * a.x_stride + a.x_off + r];
        T.re -= x2.re;
        T.im -= x2.im;
    }
    acc_update(a, oi, xv, T);
    a.out[oi] = T;
}

template <int R, int CPW, int NW, int JC, int PIPE = 0>
__global__ __launch_bounds__(NW * 64) void matvec_cols_kernel(MatvecArgs a) {
    cols_body<R, CPW, NW, JC, PIPE>(a, blockIdx.x);
}

// Grouped form: one launch covers many systems; workgroup -> (descriptor, column block)
// through a task table, and every descriptor's k is overridden by the launch's k.
template <int R, int CPW, int NW, int JC, int PIPE = 0>
__global__ __launch_bounds__(NW * 64) void matvec_cols_grouped_kernel(const MatvecArgs* __restrict__ args,
                                                                      const int* __restrict__ task_desc,
                                                                      const int* __restrict__ task_blk, int k) {
    MatvecArgs a = args[task_desc[blockIdx.x]];
    a.k = k;
    cols_body<R, CPW, NW, JC, PIPE>(a, task_blk[blockIdx.x]);
}

// Two-kernel split-K: phase A writes one partial sum per (row split, column)
// with plain stores; phase B (next launch, so the kernel boundary orders the
// stores) adds the partials in split order and applies the epilogue.
template <int R, int CPW, int U>
__global__ __launch_bounds__(256) void matvec_partial_kernel(MatvecArgs a) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i_base = blockIdx.x * (MV_WAVES * CPW) + wave * CPW;
    const int s = blockIdx.y;
    const int rows_per = (((a.N + a.split - 1) / a.split) + 63) & ~63;
    int j_begin = s * rows_per;
    int j_end = min(a.N, j_begin + rows_per);
    if (a.tri == TRI_LOWER) j_begin = max(j_begin, (i_base / 64) * 64);
    if (a.tri == TRI_UPPER) j_end = min(j_end, i_base + CPW);
    double yr[CPW][R], yi[CPW][R];
#pragma unroll
    for (int c = 0; c < CPW; ++c)
#pragma unroll
        for (int r = 0; r < R; ++r) {
            yr[c][r] = 0.0;
            yi[c][r] = 0.0;
        }
    const long ld = a.ld;
    int ic[CPW];
#pragma unroll
    for (int c = 0; c < CPW; ++c) ic[c] = min(i_base + c, a.N - 1);
    for (int j0 = j_begin; j0 < j_end; j0 += 64 * U) {
        double m[U][CPW];
        cplx xv[U][R];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = j0 + u * 64 + lane;
            const bool ok = j < j_end;
            const int jj = ok ? j : j_begin;
#pragma unroll
            for (int c = 0; c < CPW; ++c) {
                const double v = a.M[jj + ld * ic[c]];
                const int i = i_base + c;
                bool keep = ok && i < a.N;
                if (a.tri == TRI_LOWER) keep = keep && (jj >= i);
                if (a.tri == TRI_UPPER) keep = keep && (jj <= i);
                m[u][c] = keep ? v : 0.0;
            }
            const cplx* xp = a.x + (long)jj * a.x_stride + a.x_off;
#pragma unroll
            for (int r = 0; r < R; ++r) xv[u][r] = xp[r];
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int c = 0; c < CPW; ++c)
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    yr[c][r] = fma(m[u][c], xv[u][r].re, yr[c][r]);
                    yi[c][r] = fma(m[u][c], xv[u][r].im, yi[c][r]);
                }
    }
    double my_re = 0.0, my_im = 0.0;
#pragma unroll
    for (int c = 0; c < CPW; ++c)
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const double sr = wave_sum(yr[c][r]);
            const double si = wave_sum(yi[c][r]);
            if (lane == c * R + r) {
                my_re = sr;
                my_im = si;
            }
        }
    if (lane >= CPW * R) return;
    const int c = lane / R, r = lane % R;
    const int i = i_base + c;
    if (i < a.N) a.partial[((long)s * a.N + i) * R + r] = cplx{my_re, my_im};
}

__global__ void matvec_reduce_epilogue_kernel(MatvecArgs a) {
    const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (e >= (long)a.N * a.R) return;
    const int i = (int)(e / a.R), r = (int)(e % a.R);
    cplx y{0.0, 0.0};
    for (int s = 0; s < a.split; ++s) {
        const cplx p = a.partial[((long)s * a.N + i) * a.R + r];
        y.re += p.re;
        y.im += p.im;
    }
    const long oi = (long)i * a.out_stride + a.out_off + r;
    if (a.epi == EPI_PLAIN) {
        a.out[oi] = y;
        return;
    }
    const cplx xv = a.x[(long)i * a.x_stride + a.x_off + r];
    cplx T{a.alpha * y.re + a.beta * xv.re, a.alpha * y.im + a.beta * xv.im};
    if (a.epi == EPI_CHEB) {
        const cplx x2 = a.x2[(long)i * a.x_stride + a.x_off + r];
        T.re -= x2.re;
        T.im -= x2.im;
    }
    acc_update(a, oi, xv, T);
    a.out[oi] = T;
}

// Multi-RHS term as a GEMM on matrix cores: Y^T (C x N) = X^T (C x N) * M
// with C = 2R real columns (the interleaved complex state read as doubles).
// 64 x 64 output tiles (real columns x orbitals), 4 waves in 2 x 2, each wave
// 2 x 2 MFMA 16 x 16 blocks; 16-deep LDS stages; the Chebyshev epilogue runs
// from an LDS copy of the tile so that re/im of one element meet in a thread.
// PREC 0: FP64 MFMA 16x16x4 (M in FP64); PREC 1: FP32 MFMA 16x16x4 with FP32
// accumulation over the whole k range (M32 in FP32, X rounded to FP32 when
// staged); the recurrence, coefficients and accumulators stay FP64.
template <int PREC>
__global__ __launch_bounds__(256) void cheb_gemm_kernel(MatvecArgs a) {
    constexpr int BK = 16, PD = 65;
    __shared__ double As[BK][PD];  // X^T tile: [k = j][m = real column]
    __shared__ double Bs[BK][PD];  // M tile:   [k = j][n = orbital i]
    __shared__ double Ys[64][PD];  // result tile [real column][orbital]
    const int C = 2 * a.R;
    const int c0 = blockIdx.x * 64, i0 = blockIdx.y * 64;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    const double* X = reinterpret_cast<const double*>(a.x);
    const long xs = 2L * a.x_stride;  // doubles per state row
    const long xo = 2L * a.x_off;
    v4d acc64[2][2];
    v4f acc32[2][2];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            acc64[p][q] = (v4d){0.0, 0.0, 0.0, 0.0};
            acc32[p][q] = (v4f){0.f, 0.f, 0.f, 0.f};
        }
    for (int k0 = 0; k0 < a.N; k0 += BK) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int mm = t & 63, kk = (t >> 6) + 4 * r;
            const int gj = k0 + kk, gc = c0 + mm;
            double v = 0.0;
            if (gj < a.N && gc < C) v = X[gj * xs + xo + gc];
            As[kk][mm] = v;
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int kk = t & 15, nn = (t >> 4) + 16 * r;
            const int gj = k0 + kk, gi = i0 + nn;
            double v = 0.0;
            if (gj < a.N && gi < a.N) {
                if (PREC == 0) v = a.M[gj + (long)a.ld * gi];
                else v = (double)a.M32[gj + (long)a.ld * gi];
            }
            Bs[kk][nn] = v;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < BK; ks += 4) {
            double av[2], bv[2];
#pragma unroll
            for (int p = 0; p < 2; ++p) av[p] = As[ks + (lane >> 4)][wm + p * 16 + (lane & 15)];
#pragma unroll
            for (int q = 0; q < 2; ++q) bv[q] = Bs[ks + (lane >> 4)][wn + q * 16 + (lane & 15)];
#pragma unroll
            for (int p = 0; p < 2; ++p)
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    if (PREC == 0)
                        acc64[p][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[p], bv[q], acc64[p][q], 0, 0, 0);
                    else
                        acc32[p][q] = __builtin_amdgcn_mfma_f32_16x16x4f32((float)av[p], (float)bv[q], acc32[p][q],
                                                                            0, 0, 0);
                }
        }
        __syncthreads();
    }
    // accumulator -> LDS tile.  f64 layout: value v of lane l is D[l/16 + 4v][l%16];
    // f32 layout: D[4*(l/16) + v][l%16].
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = PREC == 0 ? (lane >> 4) + 4 * v : 4 * (lane >> 4) + v;
                Ys[wm + p * 16 + row][wn + q * 16 + (lane & 15)] =
                    PREC == 0 ? acc64[p][q][v] : (double)acc32[p][q][v];
            }
    __syncthreads();
    // epilogue over (orbital, complex column) pairs of the tile
    const int rc = min(32, a.R - c0 / 2);  // complex columns in this tile
    for (int e = t; e < 64 * rc; e += 256) {
        const int ii = e / rc, rr = e % rc;
        const int i = i0 + ii, r = c0 / 2 + rr;
        if (i >= a.N) continue;
        const cplx y{Ys[2 * rr][ii], Ys[2 * rr + 1][ii]};
        const long oi = (long)i * a.out_stride + a.out_off + r;
        if (a.epi == EPI_PLAIN) {
            a.out[oi] = y;
            continue;
        }
        const cplx xv = a.x[(long)i * a.x_stride + a.x_off + r];
        cplx T{a.alpha * y.re + a.beta * xv.re, a.alpha * y.im + a.beta * xv.im};
        if (a.epi == EPI_CHEB) {
            const cplx x2 = a.x2[(long)i * a.x_stride + a.x_off + r];
            T.re -= x2.re;
            T.im -= x2.im;
        }
        acc_update(a, oi, xv, T);
        a.out[oi] = T;
    }
}

// Pipelined multi-RHS term GEMM: 64 real columns x BN orbitals per workgroup,
// 32-deep k stages double-buffered in LDS with the next stage prefetched into
// registers while the current one feeds the matrix cores.  BN = 32: waves in
// 2 (columns) x 2 (orbitals), two 16x16 accumulators each; BN = 16: four waves
// of 16 columns.  LDS tiles hold FP64 (PREC 0) or FP32 (PREC 1) operands.
template <int PREC, int BN>
__global__ __launch_bounds__(256) void cheb_gemm2_kernel(MatvecArgs a) {
    using T = typename std::conditional<PREC == 0, double, float>::type;
    constexpr int BM = 64, BK = 32;
    constexpr int XPT = BK * BM / 256;  // X elements per thread per stage
    constexpr int HPT = BK * BN / 256;  // H elements per thread per stage
    constexpr int NACC = BN == 32 ? 2 : 1;
    __shared__ T As[2][BK][BM + 1];
    __shared__ T Bs[2][BK][BN + 1];
    __shared__ double Ys[BM][BN + 1];
    const int C = 2 * a.R;
    const int c0 = blockIdx.x * BM, i0 = blockIdx.y * BN;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wc = BN == 32 ? (w >> 1) * 32 : w * 16;
    const int wi = BN == 32 ? (w & 1) * 16 : 0;
    const double* X = reinterpret_cast<const double*>(a.x);
    const long xs = 2L * a.x_stride, xo = 2L * a.x_off;
    T xr[XPT], hr[HPT];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int q = 0; q < XPT; ++q) {
            const int e = t + 256 * q, kk = e / BM, cc = e % BM;
            const int gj = k0 + kk, gc = c0 + cc;
            double v = 0.0;
            if (gj < a.N && gc < C) v = X[gj * xs + xo + gc];
            xr[q] = (T)v;
        }
#pragma unroll
        for (int q = 0; q < HPT; ++q) {
            const int e = t + 256 * q, ii = e / BK, kk = e % BK;
            const int gj = k0 + kk, gi = i0 + ii;
            T v = (T)0;
            if (gj < a.N && gi < a.N) {
                if constexpr (PREC == 0) v = a.M[gj + (long)a.ld * gi];
                else v = a.M32[gj + (long)a.ld * gi];
            }
            hr[q] = v;
        }
    };
    auto stash = [&](int buf) {
#pragma unroll
        for (int q = 0; q < XPT; ++q) {
            const int e = t + 256 * q;
            As[buf][e / BM][e % BM] = xr[q];
        }
#pragma unroll
        for (int q = 0; q < HPT; ++q) {
            const int e = t + 256 * q;
            Bs[buf][e % BK][e / BK] = hr[q];
        }
    };
    v4d acc64[NACC];
    v4f acc32[NACC];
#pragma unroll
    for (int p = 0; p < NACC; ++p) {
        acc64[p] = (v4d){0.0, 0.0, 0.0, 0.0};
        acc32[p] = (v4f){0.f, 0.f, 0.f, 0.f};
    }
    fetch(0);
    stash(0);
    __syncthreads();
    int buf = 0;
    for (int k0 = 0; k0 < a.N; k0 += BK) {
        const bool more = k0 + BK < a.N;
        if (more) fetch(k0 + BK);
#pragma unroll
        for (int ks = 0; ks < BK; ks += 4) {
            const T bv = Bs[buf][ks + (lane >> 4)][wi + (lane & 15)];
#pragma unroll
            for (int p = 0; p < NACC; ++p) {
                const T av = As[buf][ks + (lane >> 4)][wc + p * 16 + (lane & 15)];
                if constexpr (PREC == 0)
                    acc64[p] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc64[p], 0, 0, 0);
                else
                    acc32[p] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc32[p], 0, 0, 0);
            }
        }
        if (more) stash(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
#pragma unroll
    for (int p = 0; p < NACC; ++p)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = PREC == 0 ? (lane >> 4) + 4 * v : 4 * (lane >> 4) + v;
            Ys[wc + p * 16 + row][wi + (lane & 15)] = PREC == 0 ? acc64[p][v] : (double)acc32[p][v];
        }
    __syncthreads();
    const int rc = min(BM / 2, a.R - c0 / 2);
    for (int e = t; e < BN * rc; e += 256) {
        const int ii = e / rc, rr = e % rc;
        const int i = i0 + ii, r = c0 / 2 + rr;
        if (i >= a.N) continue;
        const cplx y{Ys[2 * rr][ii], Ys[2 * rr + 1][ii]};
        const long oi = (long)i * a.out_stride + a.out_off + r;
        if (a.epi == EPI_PLAIN) {
            a.out[oi] = y;
            continue;
        }
        const cplx xv = a.x[(long)i * a.x_stride + a.x_off + r];
        cplx Tk{a.alpha * y.re + a.beta * xv.re, a.alpha * y.im + a.beta * xv.im};
        if (a.epi == EPI_CHEB) {
            const cplx x2 = a.x2[(long)i * a.x_stride + a.x_off + r];
            Tk.re -= x2.re;
            Tk.im -= x2.im;
        }
        acc_update(a, oi, xv, Tk);
        a.out[oi] = Tk;
    }
}

// bf16 (round to nearest even) of an FP32 value, as the bits a matrix core reads
__host__ __device__ inline unsigned short bf16_rne(float f) {
    unsigned u;
    __builtin_memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

// Chebyshev epilogue for output row i, column r with y = (H x)_ir:
//   EPI_PLAIN:       out = y
//   EPI_CHEB_FIRST:  T = alpha y + beta x; acc = a_0 x + a_1 T; out = T
//   EPI_CHEB:        T = alpha y + beta x - x2; acc += a_k T; out = T
// (a null acc: the recurrence of a multi-step series, out = T only)
__device__ __forceinline__ void term_epilogue(const MatvecArgs& a, int i, int r, cplx y) {
    const long oi = (long)i * a.out_stride + a.out_off + r;
    if (a.epi == EPI_PLAIN) {
        a.out[oi] = y;
        return;
    }
    const cplx xv = a.x[(long)i * a.x_stride + a.x_off + r];
    cplx T{a.alpha * y.re + a.beta * xv.re, a.alpha * y.im + a.beta * xv.im};
    if (a.epi == EPI_CHEB) {
        const cplx x2 = a.x2[(long)i * a.x_stride + a.x_off + r];
        T.re -= x2.re;
        T.im -= x2.im;
    }
    acc_update(a, oi, xv, T);
    a.out[oi] = T;
    if (a.out32) {  // the next term's FP32 operand: the value its GEMM would round on load
        a.out32[2 * oi] = (float)T.re;
        a.out32[2 * oi + 1] = (float)T.im;
    }
    if (a.outb16) {  // and its bf16 operand for the residual's products (precision mode 3)
        a.outb16[2 * oi] = bf16_rne((float)T.re);
        a.outb16[2 * oi + 1] = bf16_rne((float)T.im);
    }
}

// The epilogue with its operands loaded ahead (they do not depend on y).
struct EpiOperands {
    cplx xv, x2, acc, ck;
};

__device__ __forceinline__ EpiOperands epilogue_operands(const MatvecArgs& a, int i, int r) {
    EpiOperands o{};
    if (a.epi == EPI_PLAIN) return o;
    const long oi = (long)i * a.out_stride + a.out_off + r;
    o.xv = a.x[(long)i * a.x_stride + a.x_off + r];
    if (a.epi == EPI_CHEB) {
        o.x2 = a.x2[(long)i * a.x_stride + a.x_off + r];
        if (a.acc) {
            o.acc = a.acc[oi];
            o.ck = a.coef[a.k];
        }
    }
    return o;
}

// returns what it writes to out (T_k, or y for EPI_PLAIN)
__device__ __forceinline__ cplx term_epilogue_pre(const MatvecArgs& a, int i, int r, cplx y, const EpiOperands& o) {
    const long oi = (long)i * a.out_stride + a.out_off + r;
    if (a.epi == EPI_PLAIN) {
        a.out[oi] = y;
        return y;
    }
    cplx T{a.alpha * y.re + a.beta * o.xv.re, a.alpha * y.im + a.beta * o.xv.im};
    if (a.epi == EPI_CHEB) {
        T.re -= o.x2.re;
        T.im -= o.x2.im;
    }
    if (!a.acc) {
    } else if (a.epi == EPI_CHEB) {
        const cplx add = cmul(o.ck, T);
        a.acc[oi] = cplx{o.acc.re + add.re, o.acc.im + add.im};
    } else {
        const cplx t0 = cmul(a.coef[0], o.xv), t1 = cmul(a.coef[1], T);
        a.acc[oi] = cplx{t0.re + t1.re, t0.im + t1.im};
    }
    a.out[oi] = T;
    return T;
}

// ------------------------------------------------------------------ multi-step series: panels
// Accumulator q of a multi-step series takes sum_k coefm[q][k] T_k.  The
// recurrence writes T_k (k >= 1) to ring slot (k - 1) % PANEL and nothing
// else; after every PANEL terms (and after the last) one pass adds the
// panel's terms k0..k1 to the accumulators, in term order, reading each T_k
// once for all accumulators (T_0 = phi, in the first panel, which starts the
// accumulators from zero).  Accumulator q takes terms up to kq[q]; the last
// accumulator's kq is the member's series length.
// The first accumulator taking any of terms k0.. (q_lo) and the panel's last
// term for this member (kend <= k1): kq is non-decreasing, so q_lo counts the
// accumulators that end before k0.  The table is read in one round of
// independent loads.
__device__ __forceinline__ void panel_range(const PanelArgs& p, int k0, int k1, int& q_lo, int& kend) {
    int kv[MULTI_MAX];
#pragma unroll
    for (int u = 0; u < MULTI_MAX; ++u) kv[u] = u < p.m ? p.kq[u] : 0;
    q_lo = 0;
    kend = k1;
#pragma unroll
    for (int u = 0; u < MULTI_MAX; ++u) {
        if (u < p.m - 1 && kv[u] < k0) ++q_lo;
        if (u == p.m - 1) kend = min(k1, kv[u]);
    }
}

// the panel's coefficients: Cs[u * (PANEL + 1) + j] = coefm[q_lo + u][k0 + j]
// (zero past the member's accumulators or its last term), one load round for
// the workgroup; the caller synchronises before reading them
constexpr int PANEL_CS = MULTI_MAX * (PANEL + 1);
__device__ __forceinline__ void panel_coefs(const PanelArgs& p, int q_lo, int k0, int kend, cplx* Cs) {
    for (int x = threadIdx.x; x < PANEL_CS; x += blockDim.x) {
        const int u = x / (PANEL + 1), j = x % (PANEL + 1);
        const int q = q_lo + u;
        Cs[x] = (q < p.m && k0 + j <= kend) ? p.coefm[(long)q * p.coefm_stride + k0 + j] : cplx{0.0, 0.0};
    }
}

__device__ __forceinline__ void panel_body(const PanelArgs& p, int k0, int k1, long e) {
    __shared__ cplx Cs[PANEL_CS];
    int q, kend;
    panel_range(p, k0, k1, q, kend);
    panel_coefs(p, q, k0, kend, Cs);
    __syncthreads();
    if (e >= p.n || kend < k0) return;
    cplx T[PANEL + 1];
#pragma unroll
    for (int j = 0; j <= PANEL; ++j)
        if (k0 + j <= kend) {
            const int k = k0 + j;
            T[j] = k == 0 ? p.phi[e] : p.ring[(long)((k - 1) % PANEL) * p.ring_stride + e];
        }
    // every accumulator is loaded before any is stored, so that the loads
    // overlap (the stores could alias them as far as the compiler knows)
    cplx A[MULTI_MAX];
#pragma unroll
    for (int u = 0; u < MULTI_MAX; ++u)
        if (q + u < p.m) A[u] = k0 == 0 ? cplx{0.0, 0.0} : p.accm[(long)(q + u) * p.accm_stride + e];
#pragma unroll
    for (int u = 0; u < MULTI_MAX; ++u)
        if (q + u < p.m) {
            cplx acc = A[u];
#pragma unroll
            for (int j = 0; j <= PANEL; ++j)
                if (k0 + j <= kend) {
                    const cplx add = cmul(Cs[u * (PANEL + 1) + j], T[j]);
                    acc.re += add.re;
                    acc.im += add.im;
                }
            p.accm[(long)(q + u) * p.accm_stride + e] = acc;
        }
}

__global__ __launch_bounds__(256) void multi_panel_kernel(PanelArgs p, int k0, int k1) {
    panel_body(p, k0, k1, blockIdx.x * 256L + threadIdx.x);
}

// batch members: task i is (member task_desc[i], 256-element block task_blk[i])
__global__ __launch_bounds__(256) void multi_panel_grouped_kernel(const PanelArgs* __restrict__ args,
                                                                  const int* __restrict__ task_desc,
                                                                  const int* __restrict__ task_blk, int k0, int k1) {
    const PanelArgs p = args[task_desc[blockIdx.x]];
    panel_body(p, k0, k1, task_blk[blockIdx.x] * 256L + threadIdx.x);
}

void launch_multi_panel(const PanelArgs& p, int k0, int k1, hipStream_t st) {
    const unsigned blocks = (unsigned)((p.n + 255) / 256);
    if (blocks) hipLaunchKernelGGL(multi_panel_kernel, dim3(blocks), dim3(256), 0, st, p, k0, k1);
}

void launch_multi_panel_grouped(const PanelArgs* d_args, const int* task_desc, const int* task_blk, int ntask,
                                int k0, int k1, hipStream_t st) {
    if (ntask > 0)
        hipLaunchKernelGGL(multi_panel_grouped_kernel, dim3(ntask), dim3(256), 0, st, d_args, task_desc, task_blk,
                           k0, k1);
}

// Split-K multi-RHS term GEMM: Y = H X with H N x N (FP64, or its FP32 copy
// for PREC 1) and X the 2R real columns of the RHS.  A workgroup computes a
// 128 (orbitals) x 128 (real columns) tile over one K slice; 4 waves in 2 x 2,
// each a 64 x 64 quadrant of 4 x 4 MFMA 16x16x4 blocks.  16-deep k stages are
// double-buffered in LDS and prefetched into registers.  The MFMA runs with
// the operands swapped (D^T = X^T H^T), so a lane's results lie in 16
// consecutive orbitals of one column.  Slice z writes its FP64 partial
// product to P[z][c][i]; gemm3_reduce_kernel adds the slices in order and runs
// the Chebyshev epilogue.  The FP32 path accumulates each K slice in FP32 and
// adds the slices in FP64.
constexpr int G3_T = 128;
// k-stage depth: 16 (FP64) / 32 (FP32) keeps the double-buffered tiles at 74 KB of
// LDS, so two workgroups share a CU and hide each other's barrier and load waits
// PREC: 0 = FP64, 1 = FP32 products and slice sums (mixed), 2 = FP32-stored H
// widened to FP64 on load, FP64 products (the arithmetic of PREC 0)
template <int PREC> struct G3 { static constexpr int BK = PREC == 1 ? 32 : 16; };
// row pad: successive k-rows of an operand read start 32 banks apart (FP64) or
// 16 banks apart (FP32), so the four rows a wave reads do not collide
constexpr int G3_PAD = 16;

int gemm3_splits(int N, int R) {
    const int tiles = ((N + G3_T - 1) / G3_T) * ((2 * R + G3_T - 1) / G3_T);
    int S = (512 + tiles - 1) / tiles;  // two workgroups per CU
    const int kmax = (N + 255) / 256;  // at least 256 k per slice
    if (S > kmax) S = kmax;
    return S < 1 ? 1 : S;
}

long gemm3_partial_doubles(int N, int ld, int R) { return (long)gemm3_splits(N, R) * ld * 2 * R; }

template <int PREC>
__global__ __launch_bounds__(256) void gemm3_kernel(MatvecArgs a, double* __restrict__ P, int KS) {
    constexpr bool F32 = PREC == 1;
    using T = typename std::conditional<F32, float, double>::type;
    constexpr int G3_BK = G3<PREC>::BK;
    __shared__ T As[2][G3_BK][G3_T + G3_PAD];
    __shared__ T Bs[2][G3_BK][G3_T + G3_PAD];
    const int N = a.N, C = 2 * a.R;
    const int i0 = blockIdx.x * G3_T, c0 = blockIdx.y * G3_T, z = blockIdx.z;
    const int klo = z * KS, khi = min(N, klo + KS);
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wi = (w >> 1) * 64, wc = (w & 1) * 64;
    const double* X = reinterpret_cast<const double*>(a.x);
    const float* X32 = F32 ? a.x32 : nullptr;  // FP32 copy of x when the caller keeps one
    const long xs = 2L * a.x_stride, xo = 2L * a.x_off;
    using acc_t = typename std::conditional<F32, v4f, v4d>::type;
    acc_t acc[4][4];
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[p][q][v] = (T)0;
    constexpr int LPT = G3_BK * G3_T / 256;  // elements of each tile per thread per stage
    T ra[LPT], rb[LPT];
    // H tile: rows i0..i0+127 (lane-contiguous), k0..k0+BK-1; X tile: k0..k0+BK-1, columns c0..c0+127
    auto gload = [&](int k0) {
#pragma unroll
        for (int r = 0; r < LPT; ++r) {
            const int ii = t & 127, kk = (t >> 7) + 2 * r;
            const int gi = i0 + ii, gk = k0 + kk;
            T v = (T)0;
            if (gi < N && gk < khi) {
                if constexpr (PREC == 0) v = a.M[gi + (long)a.ld * gk];
                else v = (T)a.M32[gi + (long)a.ld * gk];
            }
            ra[r] = v;
        }
#pragma unroll
        for (int r = 0; r < LPT; ++r) {
            const int cc = t & 127, kk = (t >> 7) + 2 * r;
            const int gc = c0 + cc, gk = k0 + kk;
            T v = (T)0;
            if (gc < C && gk < khi) v = X32 ? (T)X32[gk * xs + xo + gc] : (T)X[gk * xs + xo + gc];
            rb[r] = v;
        }
    };
    auto swrite = [&](int s) {
#pragma unroll
        for (int r = 0; r < LPT; ++r) {
            const int kk = (t >> 7) + 2 * r;
            As[s][kk][t & 127] = ra[r];
            Bs[s][kk][t & 127] = rb[r];
        }
    };
    int s = 0;
    if (klo < khi) {
        gload(klo);
        swrite(0);
    }
    __syncthreads();
    for (int k0 = klo; k0 < khi; k0 += G3_BK) {
        const bool more = k0 + G3_BK < khi;
        if (more) gload(k0 + G3_BK);
#pragma unroll
        for (int ks = 0; ks < G3_BK; ks += 4) {
            T av[4], bv[4];
#pragma unroll
            for (int p = 0; p < 4; ++p) av[p] = As[s][ks + (lane >> 4)][wi + p * 16 + (lane & 15)];
#pragma unroll
            for (int q = 0; q < 4; ++q) bv[q] = Bs[s][ks + (lane >> 4)][wc + q * 16 + (lane & 15)];
#pragma unroll
            for (int p = 0; p < 4; ++p)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if constexpr (!F32)
                        acc[p][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[q], av[p], acc[p][q], 0, 0, 0);
                    else
                        acc[p][q] = __builtin_amdgcn_mfma_f32_16x16x4f32(bv[q], av[p], acc[p][q], 0, 0, 0);
                }
        }
        if (more) swrite(s ^ 1);
        __syncthreads();
        s ^= 1;
    }
    // D^T layout: f64 lane l, value v -> column (l >> 4) + 4v, orbital l & 15;
    //             f32 lane l, value v -> column 4 (l >> 4) + v, orbital l & 15.
    // PREC 1 stores its FP32 slice sums as FP32 (exact, half the bytes).
    using part_t = typename std::conditional<PREC == 1, float, double>::type;
    part_t* Pz = reinterpret_cast<part_t*>(P) + (long)z * C * a.ld;
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int gi = i0 + wi + p * 16 + (lane & 15);
                const int gc = c0 + wc + q * 16 + (!F32 ? (lane >> 4) + 4 * v : 4 * (lane >> 4) + v);
                if (gi < N && gc < C) Pz[(long)gc * a.ld + gi] = (part_t)acc[p][q][v];
            }
}

// slices added in order, in FP64 (PT: how the slices were stored)
template <typename PT>
__global__ __launch_bounds__(256) void gemm3_reduce_kernel(MatvecArgs a, const PT* __restrict__ P, int S) {
    const long n = (long)a.N * a.R;
    const long e = blockIdx.x * 256L + threadIdx.x;
    if (e >= n) return;
    const int r = (int)(e / a.N), i = (int)(e % a.N);
    const long C = 2L * a.R;
    cplx y{0.0, 0.0};
    for (int z = 0; z < S; ++z) {
        const PT* Pz = P + (long)z * C * a.ld;
        y.re += (double)Pz[(2L * r) * a.ld + i];
        y.im += (double)Pz[(2L * r + 1) * a.ld + i];
    }
    term_epilogue(a, i, r, y);
}

// Slices of the library GEMMs of the mixed modes, added in FP64 in slice order:
// the FP32 products Phi[z] and (mode 3) the residual's products Plo[z], both
// laid out as gemm3's slices ([z][c][i], ld rows)
__global__ __launch_bounds__(256) void lib_reduce_kernel(MatvecArgs a, const float* __restrict__ Phi,
                                                          const float* __restrict__ Plo, int S) {
    const long n = (long)a.N * a.R;
    const long e = blockIdx.x * 256L + threadIdx.x;
    if (e >= n) return;
    const int r = (int)(e / a.N), i = (int)(e % a.N);
    const long C = 2L * a.R;
    cplx y{0.0, 0.0};
    for (int z = 0; z < S; ++z) {
        const long o = (long)z * C * a.ld;
        y.re += (double)Phi[o + (2L * r) * a.ld + i];
        y.im += (double)Phi[o + (2L * r + 1) * a.ld + i];
    }
    if (Plo) {
        cplx lo{0.0, 0.0};
        for (int z = 0; z < S; ++z) {
            const long o = (long)z * C * a.ld;
            lo.re += (double)Plo[o + (2L * r) * a.ld + i];
            lo.im += (double)Plo[o + (2L * r + 1) * a.ld + i];
        }
        y.re += lo.re;
        y.im += lo.im;
    }
    term_epilogue(a, i, r, y);
}

void launch_lib_reduce(const MatvecArgs& a, const float* Phi, const float* Plo, int S, hipStream_t st) {
    const long n = (long)a.N * a.R;
    hipLaunchKernelGGL(lib_reduce_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, Phi, Plo, S);
}

void launch_gemm3(const MatvecArgs& a, int prec, double* P, hipStream_t st) {
    const int S = gemm3_splits(a.N, a.R);
    const int BK = prec == 1 ? G3<1>::BK : G3<0>::BK;
    const int KS = (((a.N + S - 1) / S) + BK - 1) / BK * BK;
    const int Sr = (a.N + KS - 1) / KS;
    dim3 grid((a.N + G3_T - 1) / G3_T, (2 * a.R + G3_T - 1) / G3_T, Sr);
    if (prec == 0) hipLaunchKernelGGL((gemm3_kernel<0>), grid, dim3(256), 0, st, a, P, KS);
    else if (prec == 1) hipLaunchKernelGGL((gemm3_kernel<1>), grid, dim3(256), 0, st, a, P, KS);
    else hipLaunchKernelGGL((gemm3_kernel<2>), grid, dim3(256), 0, st, a, P, KS);
    const long n = (long)a.N * a.R;
    const dim3 rg((unsigned)((n + 255) / 256));
    if (prec == 1)
        hipLaunchKernelGGL((gemm3_reduce_kernel<float>), rg, dim3(256), 0, st, a, reinterpret_cast<const float*>(P), Sr);
    else
        hipLaunchKernelGGL((gemm3_reduce_kernel<double>), rg, dim3(256), 0, st, a, (const double*)P, Sr);
}

static int g_gemm_bn = 0;  // 0: choose by size; 16 or 32 forces the tile width
void set_gemm_tile(int bn) { g_gemm_bn = bn; }

void launch_cheb_gemm(const MatvecArgs& a, int prec, hipStream_t st) {
    const int C = 2 * a.R;
    const int ctiles = (C + 63) / 64;
    int bn = g_gemm_bn;
    if (bn == 0) bn = ctiles * ((a.N + 31) / 32) >= 256 ? 32 : 16;
    if (bn == 32) {
        dim3 grid(ctiles, (a.N + 31) / 32);
        if (prec == 0) hipLaunchKernelGGL((cheb_gemm2_kernel<0, 32>), grid, dim3(256), 0, st, a);
        else hipLaunchKernelGGL((cheb_gemm2_kernel<1, 32>), grid, dim3(256), 0, st, a);
    } else {
        dim3 grid(ctiles, (a.N + 15) / 16);
        if (prec == 0) hipLaunchKernelGGL((cheb_gemm2_kernel<0, 16>), grid, dim3(256), 0, st, a);
        else hipLaunchKernelGGL((cheb_gemm2_kernel<1, 16>), grid, dim3(256), 0, st, a);
    }
}

int matvec_split_for(int N, int R) {
    const int cpw = R <= 4 ? 4 : (R <= 8 ? 2 : 1);
    const int colblocks = (N + MV_WAVES * cpw - 1) / (MV_WAVES * cpw);
    int split = (1024 + colblocks - 1) / colblocks;
    const int maxsplit = (N + 255) / 256;
    if (split > maxsplit) split = maxsplit;
    if (split < 1) split = 1;
    const int rows_per = (((N + split - 1) / split) + 63) & ~63;
    split = (N + rows_per - 1) / rows_per;
    return split;
}

template <int R, int CPW>
static void launch_mv_split(const MatvecArgs& a, hipStream_t st) {
    const int colblocks = (a.N + MV_WAVES * CPW - 1) / (MV_WAVES * CPW);
    hipLaunchKernelGGL((matvec_kernel<R, CPW>), dim3(colblocks, a.split), dim3(256), 0, st, a);
}

template <int R, int CPW, int U>
static void launch_mv_twophase(const MatvecArgs& a, int split, hipStream_t st) {
    MatvecArgs b = a;
    b.split = split;
    const int colblocks = (a.N + MV_WAVES * CPW - 1) / (MV_WAVES * CPW);
    hipLaunchKernelGGL((matvec_partial_kernel<R, CPW, U>), dim3(colblocks, split), dim3(256), 0, st, b);
    const long n = (long)a.N * a.R;
    hipLaunchKernelGGL(matvec_reduce_epilogue_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, b);
}

template <int R, int CPW, int NW, int JC, int PIPE = 0>
static void launch_mv_cols(const MatvecArgs& a, hipStream_t st) {
    const int nc = NW * CPW;
    const int blocks = (a.N + nc - 1) / nc;
    hipLaunchKernelGGL((matvec_cols_kernel<R, CPW, NW, JC, PIPE>), dim3(blocks), dim3(NW * 64), 0, st, a);
}

int grouped_cols_per_block(int R) { return R <= 8 ? 16 : 8; }

void launch_matvec_grouped(const MatvecArgs* d_args, const int* d_task_desc, const int* d_task_blk, int ntasks, int R,
                           int k, hipStream_t st) {
    if (ntasks <= 0) return;
    switch (R) {
        case 1: hipLaunchKernelGGL((matvec_cols_grouped_kernel<1, 2, 8, 512, 1>), dim3(ntasks), dim3(512), 0, st, d_args, d_task_desc, d_task_blk, k); break;
        case 2: hipLaunchKernelGGL((matvec_cols_grouped_kernel<2, 2, 8, 512, 1>), dim3(ntasks), dim3(512), 0, st, d_args, d_task_desc, d_task_blk, k); break;
        case 4: hipLaunchKernelGGL((matvec_cols_grouped_kernel<4, 2, 8, 512, 1>), dim3(ntasks), dim3(512), 0, st, d_args, d_task_desc, d_task_blk, k); break;
        case 8: hipLaunchKernelGGL((matvec_cols_grouped_kernel<8, 2, 8, 256, 1>), dim3(ntasks), dim3(512), 0, st, d_args, d_task_desc, d_task_blk, k); break;
        case 16: hipLaunchKernelGGL((matvec_cols_grouped_kernel<16, 1, 8, 128, 1>), dim3(ntasks), dim3(512), 0, st, d_args, d_task_desc, d_task_blk, k); break;
        case 32: hipLaunchKernelGGL((matvec_cols_grouped_kernel<32, 1, 8, 64, 1>), dim3(ntasks), dim3(512), 0, st, d_args, d_task_desc, d_task_blk, k); break;
        default: break;
    }
}

// One strided complex column to another (host transfers go through a contiguous
// staging column and this copy).
__global__ __launch_bounds__(256) void strided_copy_kernel(cplx* dst, long ds, const cplx* src, long ss, int n) {
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i < n) dst[i * ds] = src[i * ss];
}

void launch_strided_copy(cplx* dst, long ds, const cplx* src, long ss, int n, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(strided_copy_kernel, dim3((n + 255) / 256), dim3(256), 0, st, dst, ds, src, ss, n);
}

// ------------------------------------------------------------------ symmetric half-storage term
// H_orth is symmetric, so a term only needs the 64 x 64 tiles (I, J) with
// J <= I: each off-diagonal tile is read once and applied twice, as H_IJ to
// x_J (rows of block I) and as H_IJ^T to x_I (rows of block J).
//
// Phase 1 (sym_tiles_kernel): a workgroup takes a segment of S consecutive
// tiles of block row I (S is picked so the grid fits the resident slots).
//  * rows: lane l of wave w holds row i0+l for columns 16w..16w+15, read as
//    coalesced 512 B column segments; x_J[j] is the same for the whole wave
//    and comes in through scalar loads, and the row sums stay in registers
//    over the segment;
//  * columns: the tile goes through LDS; lane l of wave w then holds column
//    j0+l for rows 16w..16w+15 (x_I again wave-uniform), and the four waves'
//    column sums are added in LDS.
// Column sums of tile (I, J) go to Pcol[I(I-1)/2 + J], row sums of segment s
// to Prow[s] (64 x R complex each, in a.partial).
// Phase 2 (sym_reduce_kernel): for block B, adds the segments of block row B
// and Pcol[(I, B)] for I > B in a fixed order (deterministic), then runs the
// Chebyshev epilogue.
__host__ __device__ inline int sym_segments_in_row(int I, int S) { return (I + S) / S; }

// segments of block rows 0..B-1: sum_{k=1..B} ceil(k / S) in closed form
__host__ __device__ inline int sym_segment_start(int B, int S) {
    const int q = B / S, r = B - q * S;
    return S * q * (q + 1) / 2 + r * (q + 1);
}

static int sym_segment_count(int nb, int S) { return sym_segment_start(nb, S); }

// shortest segment whose grid fits per_cu resident workgroups per CU (256 CUs):
// four for the vector kernel, three for the matrix-core kernel (its LDS)
static int sym_seg_for(int nb, int per_cu = 4) {
    int S = 1;
    while (sym_segment_count(nb, S) > per_cu * 256) ++S;
    return S;
}

// the partial buffer of the one-row-segment symmetric kernels: the segment
// count only falls as per_cu falls, so the four-per-CU grid is the largest of
// the resident grids; the fixed three-tile segments (variant 21) may have more
long sym_partial_cplx(int ld, int R) {
    const long nb = ld / TILE;
    const long segs = std::max(sym_segment_count((int)nb, sym_seg_for((int)nb, 4)), sym_segment_count((int)nb, 3));
    return (nb * (nb - 1) / 2 + segs) * TILE * R;
}

bool sym_supported(int R, int tri) { return tri == TRI_FULL && (R == 1 || R == 2 || R == 4 || R == 8); }


// x and H are read-only here and restrict-qualified, so wave-uniform x reads
// become scalar loads.
template <int R>
__global__ __launch_bounds__(256) void sym_tiles_kernel(const double* __restrict__ Mg, const cplx* __restrict__ xg,
                                                        cplx* __restrict__ part, int ld, int N, int x_stride,
                                                        int x_off, int S) {
    constexpr int HP = TILE + 1;  // padded LDS column
    constexpr int XB = 16 / R;    // x rows per scalar-load batch (4R dwords each)
    // the four waves' column sums reuse the tile's LDS once it has been read:
    // 33 KB per workgroup, four workgroups per CU
    static_assert(4 * TILE * 2 * R <= TILE * HP, "column sums fit in the tile buffer");
    __shared__ double Hs[TILE * HP];
    double* const zred = Hs;
    const int nb = ld / TILE;
    // block row of this segment: the last I with sym_segment_start(I) <= blockIdx.x
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sym_segment_start(mid, S) <= (int)blockIdx.x) lo = mid;
        else hi = mid - 1;
    }
    const int I = lo;
    const int Ja = ((int)blockIdx.x - sym_segment_start(I, S)) * S, Jb = min(I + 1, Ja + S);
    const int t = threadIdx.x, lane = t & 63;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int i0 = I * TILE, row = i0 + lane;
    // H is read over the whole padded ld x ld tile.  Its padding is
    // [0 0; 0 I] (from the identity-padded S and h, exactly), so H[i][j] = 0
    // whenever one of i, j is < N and the other >= N: rows >= N of x never
    // reach rows < N of the result, and they are read as row N - 1 (finite
    // data) instead of the unwritten padding.  Results for rows >= N are
    // discarded by phase 2.
    double hn[16];
    auto fetch = [&](int J) {
        const int j0 = J * TILE + 16 * w;
#pragma unroll
        for (int s = 0; s < 16; ++s) hn[s] = Mg[(long)(j0 + s) * ld + row];
    };
    auto xload = [&](int j, double* xr, double* xi) {  // wave-uniform row j of x
        const cplx* p = xg + (long)min(j, N - 1) * x_stride + x_off;
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const cplx v = p[q];
            xr[q] = v.re;
            xi[q] = v.im;
        }
    };
    double yr[R], yi[R];
#pragma unroll
    for (int q = 0; q < R; ++q) yr[q] = yi[q] = 0.0;
    fetch(Ja);
    for (int J = Ja; J < Jb; ++J) {
        const int j0 = J * TILE;
        __syncthreads();  // previous tile's LDS readers are done
        double hc[16];
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            hc[s] = hn[s];
            Hs[(16 * w + s) * HP + lane] = hn[s];
        }
        if (J + 1 < Jb) fetch(J + 1);
        // rows of block I: H_IJ x_J over this wave's 16 columns.  Scalar loads
        // return out of order (only a wait for all of them is possible), so
        // they are issued in batches of XB rows (~64 SGPRs), then consumed.
#pragma unroll
        for (int s0 = 0; s0 < 16; s0 += XB) {
            double xr[XB][R], xi[XB][R];
#pragma unroll
            for (int u = 0; u < XB; ++u) xload(j0 + 16 * w + s0 + u, xr[u], xi[u]);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int u = 0; u < XB; ++u)
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    yr[q] = fma(hc[s0 + u], xr[u][q], yr[q]);
                    yi[q] = fma(hc[s0 + u], xi[u][q], yi[q]);
                }
#pragma unroll
            for (int q = 0; q < R; ++q) asm volatile("" : "+v"(yr[q]), "+v"(yi[q]));
            __builtin_amdgcn_sched_barrier(0);
        }
        if (J == I) continue;  // diagonal tile: the row sums cover it
        __syncthreads();       // tile complete in LDS
        // rows of block J: H_IJ^T x_I, column j0+lane over rows 16w..16w+15
        double zr[R], zi[R];
#pragma unroll
        for (int q = 0; q < R; ++q) zr[q] = zi[q] = 0.0;
#pragma unroll
        for (int s0 = 0; s0 < 16; s0 += XB) {
            double xr[XB][R], xi[XB][R], hv[XB];
#pragma unroll
            for (int u = 0; u < XB; ++u) {
                hv[u] = Hs[lane * HP + 16 * w + s0 + u];
                xload(i0 + 16 * w + s0 + u, xr[u], xi[u]);
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int u = 0; u < XB; ++u)
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    zr[q] = fma(hv[u], xr[u][q], zr[q]);
                    zi[q] = fma(hv[u], xi[u][q], zi[q]);
                }
#pragma unroll
            for (int q = 0; q < R; ++q) asm volatile("" : "+v"(zr[q]), "+v"(zi[q]));
            __builtin_amdgcn_sched_barrier(0);
        }
        __syncthreads();  // every wave is done with the tile
#pragma unroll
        for (int q = 0; q < R; ++q) {
            zred[((w * TILE) + lane) * 2 * R + 2 * q] = zr[q];
            zred[((w * TILE) + lane) * 2 * R + 2 * q + 1] = zi[q];
        }
        __syncthreads();
        const long tI = (long)I * (I - 1) / 2 + J;
        for (int e = t; e < TILE * R; e += 256) {
            const int c = e / R, q = e % R;
            double sr = 0.0, si = 0.0;
#pragma unroll
            for (int ww = 0; ww < 4; ++ww) {
                sr += zred[(ww * TILE + c) * 2 * R + 2 * q];
                si += zred[(ww * TILE + c) * 2 * R + 2 * q + 1];
            }
            part[tI * TILE * R + e] = cplx{sr, si};
        }
    }
    // row sums: add the four waves' column groups in order
    __syncthreads();
#pragma unroll
    for (int q = 0; q < R; ++q) {
        zred[(w * TILE + lane) * 2 * R + 2 * q] = yr[q];
        zred[(w * TILE + lane) * 2 * R + 2 * q + 1] = yi[q];
    }
    __syncthreads();
    const long pcol = (long)nb * (nb - 1) / 2 * TILE * R;
    for (int e = t; e < TILE * R; e += 256) {
        const int rr = e / R, q = e % R;
        double sr = 0.0, si = 0.0;
#pragma unroll
        for (int ww = 0; ww < 4; ++ww) {
            sr += zred[(ww * TILE + rr) * 2 * R + 2 * q];
            si += zred[(ww * TILE + rr) * 2 * R + 2 * q + 1];
        }
        part[pcol + (long)blockIdx.x * TILE * R + e] = cplx{sr, si};
    }
}

// Phase 2: workgroup (B, c) owns 64 consecutive complex outputs of block B
// (rows B*64 + (c*64 + e) / R, column (c*64 + e) % R).  G groups of 64
// threads split the source list; each group issues up to SYM_MAXS loads
// before adding them (latency, not bandwidth, bounds this kernel), and the
// group sums are added in group order.  G follows the block count: a block
// has about nb sources, so G = 16 spreads the 64 sources of nb = 64 four per
// thread, and G = 4 keeps a small system's workgroup busy (G = 16 left most
// of its groups without a source).  The order of additions depends on (nb, G)
// only, so a system gives the same bits alone and in a batch.
constexpr int SYM_G = 16, SYM_MAXS = 8;
constexpr int SYM_REDUCE_THREADS = SYM_G * 64;  // one 64-thread group per source group
static_assert(SYM_REDUCE_THREADS <= 1024, "phase-2 workgroup size");
__host__ __device__ inline int sym_reduce_groups(int nb) { return nb > 32 ? 16 : 4; }

// MULTI: a multi-step series' accumulation (a.accm): the terms go to
// accumulators m_lo..m_hi-1 with their own coefficients, two at a time
// (a.m_pair = 2: c_{k-1} x + c_k T_k, where x = T_{k-1} is this term's input;
// the term before, a.m_pair = 0, adds nothing; a.m_pair = 1: c_k T_k alone, a
// series' last even term).  Group g of the workgroup owns accumulators m_lo +
// g + v G (v < QP), prefetched with their coefficients alongside the
// epilogue's operands, and updates them with x and T_k from LDS at the end, so
// that no thread updates more than QP of them.  The additions are those of
// one term at a time, in term order.
template <int R, int G, bool MULTI = false>
__device__ inline void sym_reduce_body(const MatvecArgs& a, int S, int B, int chunk) {
    __shared__ cplx red[G * 64];
    __shared__ cplx Ts[MULTI ? 64 : 1], Xs[MULTI ? 64 : 1];
    constexpr int QP = (MULTI_MAX + G - 1) / G;
    const int nb = a.ld / TILE;
    const int t = threadIdx.x, g = t >> 6, e = chunk * 64 + (t & 63);
    const int seg0 = sym_segment_start(B, S);
    const int nseg = sym_segments_in_row(B, S);
    const int L = nseg + (nb - 1 - B);   // sources: row segments, then Pcol(I, B), I > B
    const int s0 = g * L / G, s1 = (g + 1) * L / G;
    const long E = (long)TILE * R;
    const long pcol = (long)nb * (nb - 1) / 2 * E;
    const int i = B * TILE + e / R;
    const long oi = (long)i * a.out_stride + a.out_off + e % R;
    const bool valid = i < a.N;
    // the epilogue's operands do not depend on the partials: fetch them first
    EpiOperands ep{};
    if (t < 64 && valid) ep = epilogue_operands(a, i, e % R);
    const bool spread = MULTI && a.epi == EPI_CHEB && a.m_pair > 0;
    const bool pair = MULTI && a.m_pair == 2;
    cplx am[MULTI ? QP : 1], cm[MULTI ? QP : 1], cx[MULTI ? QP : 1];
    if (spread) {
#pragma unroll
        for (int v = 0; v < QP; ++v) {
            const int q = a.m_lo + g + v * G;
            if (valid && q < a.m_hi) {
                am[v] = a.accm[(long)q * a.accm_stride + oi];
                cm[v] = a.coefm[(long)q * a.coefm_stride + a.k];
                if (pair) cx[v] = a.coefm[(long)q * a.coefm_stride + a.k - 1];
            }
        }
    }
    double sr = 0.0, si = 0.0;
    for (int base = s0; base < s1; base += SYM_MAXS) {
        cplx p[SYM_MAXS];
#pragma unroll
        for (int u = 0; u < SYM_MAXS; ++u) {
            const int src = min(base + u, s1 - 1);
            const long I = B + 1 + (src - nseg);
            const long off = src < nseg ? pcol + (long)(seg0 + src) * E : (I * (I - 1) / 2 + B) * E;
            p[u] = a.partial[off + e];
        }
#pragma unroll
        for (int u = 0; u < SYM_MAXS; ++u)
            if (base + u < s1) {
                sr += p[u].re;
                si += p[u].im;
            }
    }
    red[t] = cplx{sr, si};
    __syncthreads();
    if (t < 64) {
        cplx y{0.0, 0.0};
#pragma unroll
        for (int gg = 0; gg < G; ++gg) {
            y.re += red[gg * 64 + t].re;
            y.im += red[gg * 64 + t].im;
        }
        if (valid) {
            const cplx T = term_epilogue_pre(a, i, e % R, y, ep);
            if (MULTI) {
                Ts[t] = T;
                Xs[t] = ep.xv;
                if (a.epi == EPI_CHEB_FIRST)  // one term starts every accumulator
                    for (int q = a.m_lo; q < a.m_hi; ++q) {
                        const cplx* cq = a.coefm + (long)q * a.coefm_stride;
                        const cplx t0 = cmul(cq[0], ep.xv), t1 = cmul(cq[1], T);
                        a.accm[(long)q * a.accm_stride + oi] = cplx{t0.re + t1.re, t0.im + t1.im};
                    }
            }
        }
    }
    if (spread) {
        __syncthreads();  // x and T_k of the 64 elements in LDS
        const cplx T = Ts[t & 63];
        const cplx X = pair ? Xs[t & 63] : cplx{0.0, 0.0};
#pragma unroll
        for (int v = 0; v < QP; ++v) {
            const int q = a.m_lo + g + v * G;
            if (valid && q < a.m_hi) {
                cplx acc = am[v];
                if (pair) {
                    const cplx a1 = cmul(cx[v], X);
                    acc.re += a1.re;
                    acc.im += a1.im;
                }
                const cplx add = cmul(cm[v], T);
                a.accm[(long)q * a.accm_stride + oi] = cplx{acc.re + add.re, acc.im + add.im};
            }
        }
    }
}

// MULTI: compiled apart so that the other launches keep their registers
template <int R, int G, bool MULTI>
__global__ __launch_bounds__(G * 64) void sym_reduce_kernel(MatvecArgs a, int S) {
    sym_reduce_body<R, G, MULTI>(a, S, blockIdx.x, blockIdx.y);
}

static thread_local bool t_multi_done = false;
bool take_multi_done() {
    const bool f = t_multi_done;
    t_multi_done = false;
    return f;
}

template <int R>
static void launch_sym_reduce(const MatvecArgs& a, int nb, int S, hipStream_t st) {
    // a term of a multi-step series that adds nothing (m_pair = 0) runs the
    // plain kernel; the series is still accumulated by this kernel kind
    const bool multi = a.accm != nullptr && (a.epi == EPI_CHEB_FIRST || a.m_pair > 0);
    if (sym_reduce_groups(nb) == 16) {
        if (multi) hipLaunchKernelGGL((sym_reduce_kernel<R, 16, true>), dim3(nb, R), dim3(16 * 64), 0, st, a, S);
        else hipLaunchKernelGGL((sym_reduce_kernel<R, 16, false>), dim3(nb, R), dim3(16 * 64), 0, st, a, S);
    } else {
        if (multi) hipLaunchKernelGGL((sym_reduce_kernel<R, 4, true>), dim3(nb, R), dim3(4 * 64), 0, st, a, S);
        else hipLaunchKernelGGL((sym_reduce_kernel<R, 4, false>), dim3(nb, R), dim3(4 * 64), 0, st, a, S);
    }
    if (a.accm) t_multi_done = true;
}

// Batches: task i is (system task_desc[i], block * R + chunk = task_idx[i]),
// with the system's segment length sys_S[task_desc[i]]
template <int R, int G>
__global__ __launch_bounds__(G * 64) void sym_reduce_grouped_kernel(const MatvecArgs* __restrict__ args,
                                                                    const int* __restrict__ task_desc,
                                                                    const int* __restrict__ task_idx,
                                                                    const int* __restrict__ sys_S, int k) {
    const int d = task_desc[blockIdx.x], v = task_idx[blockIdx.x];
    MatvecArgs a = args[d];
    a.k = k;
    sym_reduce_body<R, G>(a, sys_S[d], v / R, v % R);
}

template <int R>
static void launch_sym(const MatvecArgs& a, hipStream_t st) {
    const int nb = a.ld / TILE;
    const int S = sym_seg_for(nb);
    hipLaunchKernelGGL((sym_tiles_kernel<R>), dim3(sym_segment_count(nb, S)), dim3(256), 0, st, a.M, a.x,
                       a.partial, a.ld, a.N, a.x_stride, a.x_off, S);
    launch_sym_reduce<R>(a, nb, S, st);
}

// ------------------------------------------------------------------ symmetric, matrix-core tiles
// Variant 17: the segments and partial layout of the symmetric kernel above
// (so sym_reduce_kernel is its phase 2), with both tile products on FP64 MFMA
// 16x16x4.  Wave w of the workgroup owns rows 16w..16w+15 of the row product
// and columns 16w..16w+15 of the column product, each complete over the tile:
//  * rows:    Y^T[rhs][i] += sum_k X_J^T[rhs][k] H[i][k]; the B operand
//             (lane l: H[16w + (l & 15)][4c + (l >> 4)]) is the wave's own
//             coalesced load (4 columns x 128 B per instruction);
//  * columns: Z^T[rhs][j] += sum_k X_I^T[rhs][k] H[k][j]; the B operand
//             (lane l: H[4r + (l >> 4)][16w + (l & 15)]) comes from the tile in
//             LDS (column-major, 65-double columns: conflict-free both ways).
// x_J (per tile) and x_I (per segment) are staged in LDS as 64 x 2R doubles
// and read as the A operands (lane l: row 4c + (l >> 4), real column l & 15,
// clamped to 2R - 1; MFMA output rows >= 2R are discarded).  No scalar loads
// and no cross-wave reductions; partial sums are written straight from the
// accumulators (lane l, value v: real column (l >> 4) + 4v, row/column l & 15).
//
// MF = 1 (variant 18) runs the same products as FP64 MFMA 4x4x4 in 4 blocks,
// whose lanes are A: 16k + 4b + i, B: 16k + 4b + j, D: 16i + 4b + j (checked
// on the device).  Block b covers rows (or columns) 4b..4b+3 of the wave's 16,
// so the A operand of the row product is again the wave's own load
// (4b + i = l & 15) and that of the column product the same LDS read; B is x
// in groups of 4 real columns (lane l: column 4h + (l & 3)).  For R <= 4 this
// does no discarded work: ceil(2R / 4) products per k-step of 16 cycles each,
// against one 64-cycle 16x16x4 whose rows >= 2R are discarded.
template <int R, int MF, typename MT = double>
__device__ inline void symm_tiles_body(const MT* __restrict__ Mg, const double* __restrict__ xg,
                                       double* __restrict__ part, int ld, int N, int x_stride, int x_off, int S,
                                       int seg) {
    constexpr int HP = TILE + 1;
    constexpr int C2 = 2 * R;                   // real right-hand-side columns
    constexpr int XE = TILE * C2;               // doubles of one staged x block
    constexpr int XPT = (XE + 255) / 256;       // staged x doubles per thread
    __shared__ double Hs[TILE * HP];
    __shared__ double xs[XE];                   // x_J of the current tile
    __shared__ double xis[XE];                  // x_I of the segment
    const int nb = ld / TILE;
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sym_segment_start(mid, S) <= seg) lo = mid;
        else hi = mid - 1;
    }
    const int I = lo;
    const int Ja = (seg - sym_segment_start(I, S)) * S, Jb = min(I + 1, Ja + S);
    const int t = threadIdx.x, lane = t & 63;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int i0 = I * TILE;
    const int lr = lane & 15, lk = lane >> 4;
    const int xc = min(lr, C2 - 1);             // A-operand real column of this lane
    const long xs2 = 2L * x_stride, xo2 = 2L * x_off;
    // H's padding is [0 0; 0 I] and x rows >= N are read as row N - 1 (finite
    // data): rows >= N of x never reach rows < N of the result.
    auto xfetch = [&](int r0, double* v) {      // 64 x 2R block of x from row r0
#pragma unroll
        for (int u = 0; u < XPT; ++u) {
            const int e = t + 256 * u;
            if (XE % 256 == 0 || e < XE) v[u] = xg[(long)min(r0 + e / C2, N - 1) * xs2 + xo2 + e % C2];
        }
    };
    double hn[16], xn[XPT];
    auto fetch = [&](int J) {
        const int j0 = J * TILE;
#pragma unroll
        for (int c = 0; c < 16; ++c) hn[c] = (double)Mg[(long)(j0 + 4 * c + lk) * ld + i0 + 16 * w + lr];
        xfetch(j0, xn);
    };
    {
        double xi[XPT];
        xfetch(i0, xi);
#pragma unroll
        for (int u = 0; u < XPT; ++u) {
            const int e = t + 256 * u;
            if (XE % 256 == 0 || e < XE) xis[e] = xi[u];
        }
    }
    constexpr int G = (C2 + 3) / 4;             // MF = 1: groups of 4 real columns
    v4d yacc = {0.0, 0.0, 0.0, 0.0};
    double yacc4[G];
#pragma unroll
    for (int h = 0; h < G; ++h) yacc4[h] = 0.0;
    const int xq = lane & 3;
    fetch(Ja);
    for (int J = Ja; J < Jb; ++J) {
        double hc[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) hc[c] = hn[c];
        __syncthreads();  // previous tile's LDS readers are done
#pragma unroll
        for (int c = 0; c < 16; ++c) Hs[(4 * c + lk) * HP + 16 * w + lr] = hc[c];
#pragma unroll
        for (int u = 0; u < XPT; ++u) {
            const int e = t + 256 * u;
            if (XE % 256 == 0 || e < XE) xs[e] = xn[u];
        }
        if (J + 1 < Jb) fetch(J + 1);
        __syncthreads();  // tile and x_J staged
        if constexpr (MF == 0) {
#pragma unroll
            for (int c = 0; c < 16; ++c)
                yacc = __builtin_amdgcn_mfma_f64_16x16x4f64(xs[(4 * c + lk) * C2 + xc], hc[c], yacc, 0, 0, 0);
        } else {
#pragma unroll
            for (int c = 0; c < 16; ++c)
#pragma unroll
                for (int h = 0; h < G; ++h)
                    yacc4[h] = __builtin_amdgcn_mfma_f64_4x4x4f64(hc[c], xs[(4 * c + lk) * C2 + min(4 * h + xq, C2 - 1)],
                                                                  yacc4[h], 0, 0, 0);
        }
        if (J == I) continue;  // diagonal tile: the row sums cover it
        const long tI = (long)I * (I - 1) / 2 + J;
        if constexpr (MF == 0) {
            v4d zacc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int r = 0; r < 16; ++r)
                zacc = __builtin_amdgcn_mfma_f64_16x16x4f64(xis[(4 * r + lk) * C2 + xc],
                                                            Hs[(16 * w + lr) * HP + 4 * r + lk], zacc, 0, 0, 0);
            double* pz = part + 2 * (tI * TILE * R + (long)(16 * w + lr) * R);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int m = lk + 4 * v;
                if (m < C2) pz[m] = zacc[v];
            }
        } else {
            double zacc4[G];
#pragma unroll
            for (int h = 0; h < G; ++h) zacc4[h] = 0.0;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const double hv = Hs[(16 * w + lr) * HP + 4 * r + lk];
#pragma unroll
                for (int h = 0; h < G; ++h)
                    zacc4[h] = __builtin_amdgcn_mfma_f64_4x4x4f64(hv, xis[(4 * r + lk) * C2 + min(4 * h + xq, C2 - 1)],
                                                                  zacc4[h], 0, 0, 0);
            }
            // D lane l: column 16w + 4 ((l >> 2) & 3) + (l >> 4), real column 4h + (l & 3)
            double* pz = part + 2 * (tI * TILE * R + (long)(16 * w + 4 * ((lane >> 2) & 3) + lk) * R);
#pragma unroll
            for (int h = 0; h < G; ++h)
                if (4 * h + xq < C2) pz[4 * h + xq] = zacc4[h];
        }
    }
    const long pcol = (long)nb * (nb - 1) / 2 * TILE * R;
    if constexpr (MF == 0) {
        double* py = part + 2 * (pcol + (long)seg * TILE * R + (long)(16 * w + lr) * R);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int m = lk + 4 * v;
            if (m < C2) py[m] = yacc[v];
        }
    } else {
        double* py = part + 2 * (pcol + (long)seg * TILE * R + (long)(16 * w + 4 * ((lane >> 2) & 3) + lk) * R);
#pragma unroll
        for (int h = 0; h < G; ++h)
            if (4 * h + xq < C2) py[4 * h + xq] = yacc4[h];
    }
}

template <int R, int MF, typename MT = double>
__global__ __launch_bounds__(256) void symm_tiles_kernel(const MT* __restrict__ Mg, const double* __restrict__ xg,
                                                         double* __restrict__ part, int ld, int N, int x_stride,
                                                         int x_off, int S) {
    symm_tiles_body<R, MF, MT>(Mg, xg, part, ld, N, x_stride, x_off, S, blockIdx.x);
}

// Batches: task i is segment task_idx[i] of system task_desc[i] (FP64 H, 4x4x4
// products: the variant-18 arithmetic of each system, segment for segment)
template <int R>
__global__ __launch_bounds__(256) void symm_tiles_grouped_kernel(const MatvecArgs* __restrict__ args,
                                                                 const int* __restrict__ task_desc,
                                                                 const int* __restrict__ task_idx,
                                                                 const int* __restrict__ sys_S) {
    const int d = task_desc[blockIdx.x];
    const MatvecArgs& a = args[d];
    symm_tiles_body<R, 1, double>(a.M, reinterpret_cast<const double*>(a.x), reinterpret_cast<double*>(a.partial),
                                  a.ld, a.N, a.x_stride, a.x_off, sys_S[d], task_idx[blockIdx.x]);
}

// Segment length S and segment count of a system's symmetric term in a batch:
// up to 8 tiles of a block row per workgroup.  A batch has many systems to fill
// the device, so it takes long segments (row sums stay in registers across the
// segment: fewer partial sums to write and to add) where a lone system
// shortens them to fill 256 CUs.  S depends on ld only, so a system gives the
// same bits in every batch.
int sym_term_segments(int ld, int* S) {
    const int nb = ld / TILE;
    *S = nb < 8 ? nb : 8;
    return sym_segment_count(nb, *S);
}

void launch_sym_grouped(const MatvecArgs* d_args, const int* t1_desc, const int* t1_idx, int n1, const int* t2_desc,
                        const int* t2_idx, int n2, const int* sys_S, int R, int k, int reduce_groups, hipStream_t st) {
    if (n1 > 0) switch (R) {
            case 1: hipLaunchKernelGGL(symm_tiles_grouped_kernel<1>, dim3(n1), dim3(256), 0, st, d_args, t1_desc, t1_idx, sys_S); break;
            case 2: hipLaunchKernelGGL(symm_tiles_grouped_kernel<2>, dim3(n1), dim3(256), 0, st, d_args, t1_desc, t1_idx, sys_S); break;
            case 4: hipLaunchKernelGGL(symm_tiles_grouped_kernel<4>, dim3(n1), dim3(256), 0, st, d_args, t1_desc, t1_idx, sys_S); break;
            case 8: hipLaunchKernelGGL(symm_tiles_grouped_kernel<8>, dim3(n1), dim3(256), 0, st, d_args, t1_desc, t1_idx, sys_S); break;
            default: break;
        }
    if (n2 <= 0) return;
#define DH_SYM_RED(RR, GG)                                                                                 \
    hipLaunchKernelGGL((sym_reduce_grouped_kernel<RR, GG>), dim3(n2), dim3(GG * 64), 0, st, d_args, t2_desc, t2_idx, \
                       sys_S, k)
    if (reduce_groups == 16) switch (R) {
            case 1: DH_SYM_RED(1, 16); break;
            case 2: DH_SYM_RED(2, 16); break;
            case 4: DH_SYM_RED(4, 16); break;
            case 8: DH_SYM_RED(8, 16); break;
            default: break;
        }
    else switch (R) {
            case 1: DH_SYM_RED(1, 4); break;
            case 2: DH_SYM_RED(2, 4); break;
            case 4: DH_SYM_RED(4, 4); break;
            case 8: DH_SYM_RED(8, 4); break;
            default: break;
        }
#undef DH_SYM_RED
}

int sym_reduce_groups_for(int ld) { return sym_reduce_groups(ld / TILE); }

// MT = float: H's FP32 copy (a.M32), widened to FP64 on load (precision mode 2)
template <int R, int MF, typename MT = double>
static void launch_symm(const MatvecArgs& a, hipStream_t st) {
    const int nb = a.ld / TILE;
    const int S = sym_seg_for(nb, 3);
    const MT* M;
    if constexpr (std::is_same<MT, float>::value) M = a.M32;
    else M = a.M;
    hipLaunchKernelGGL((symm_tiles_kernel<R, MF, MT>), dim3(sym_segment_count(nb, S)), dim3(256), 0, st, M,
                       reinterpret_cast<const double*>(a.x), reinterpret_cast<double*>(a.partial), a.ld, a.N,
                       a.x_stride, a.x_off, S);
    launch_sym_reduce<R>(a, nb, S, st);
}

// ------------------------------------------------------------------ symmetric, column strips
// Variant 20: the segments and partial layout of variants 16-18 (so
// sym_reduce_kernel is its phase 2), with no LDS staging of the tile and no
// barrier per tile.  Wave w owns columns 16w..16w+15 of every tile of the
// segment (all 64 rows), in 16 registers (rg, ks): lane 16k + 4b + i holds
// H[16rg + 4b + i][16w + 4ks + k] (4 columns x 128 B per load instruction).
// FP64 MFMA 4x4x4 (4 blocks) lanes: A 16k + 4b + i, B 16k + 4b + j,
// D 16i + 4b + j.
//  * rows: each register is the A operand as loaded (block b = rows
//    16rg + 4b..+3, reduction over the wave's columns 16w + 4ks + k); B is
//    x_J (lane 16k + 4b + n: x_J[16w + 4ks + k][4h + n]).  The sums run over
//    the whole segment; the four waves' row sums (each over its own columns)
//    are added in LDS once per segment, in wave order;
//  * columns: a product with the identity as B moves A[i][k] of lane
//    16k + 4b + i to lane 16i + 4b + k, exactly (one product by 1, three by
//    0), which is the A operand with the rows 16rg + 4b + k as the reduction
//    index; B is x_I (lane 16k + 4b + n: x_I[16rg + 4b + k][4h + n]).  Summed
//    over rg in the accumulator and over the four blocks (lanes 4 and 8
//    apart) by two DPP row rotations, these are the complete column sums of
//    the wave's 16 columns, written straight to the tile's partial.
template <int CTRL>
__device__ inline double dpp_f64(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
constexpr int DPP_ROW_ROR4 = 0x124, DPP_ROW_ROR8 = 0x128;

template <int R>
__global__ __launch_bounds__(256) void symm3_tiles_kernel(const double* __restrict__ Mg, const double* __restrict__ xg,
                                                          double* __restrict__ part, int ld, int N, int x_stride,
                                                          int x_off, int S) {
    constexpr int C2 = 2 * R;
    constexpr int G = (C2 + 3) / 4;   // groups of 4 real columns
    __shared__ double ys[4 * TILE * C2];
    const int nb = ld / TILE;
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sym_segment_start(mid, S) <= (int)blockIdx.x) lo = mid;
        else hi = mid - 1;
    }
    const int I = lo;
    const int Ja = ((int)blockIdx.x - sym_segment_start(I, S)) * S, Jb = min(I + 1, Ja + S);
    const int t = threadIdx.x, lane = t & 63;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int i0 = I * TILE;
    const int lk = lane >> 4, lb = (lane >> 2) & 3, lq = lane & 3;
    const long xs2 = 2L * x_stride, xo2 = 2L * x_off;
    // H's padding is [0 0; 0 I] and x rows >= N are read as row N - 1 (finite
    // data): rows >= N of x never reach rows < N of the result.  Real columns
    // >= 2R are read as column 2R - 1; their MFMA outputs are not stored.
    auto xval = [&](int r, int h) { return xg[(long)min(r, N - 1) * xs2 + xo2 + min(4 * h + lq, C2 - 1)]; };
    double xi[4][G];
#pragma unroll
    for (int rg = 0; rg < 4; ++rg)
#pragma unroll
        for (int h = 0; h < G; ++h) xi[rg][h] = xval(i0 + 16 * rg + 4 * lb + lk, h);
    const double ident = lk == lq ? 1.0 : 0.0;
    double hn[16], xn[4][G];
    auto fetch = [&](int J) {
        const int j0 = J * TILE + 16 * w;
#pragma unroll
        for (int rg = 0; rg < 4; ++rg)
#pragma unroll
            for (int ks = 0; ks < 4; ++ks)
                hn[4 * rg + ks] = Mg[(long)(j0 + 4 * ks + lk) * ld + i0 + 16 * rg + 4 * lb + lq];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
#pragma unroll
            for (int h = 0; h < G; ++h) xn[ks][h] = xval(j0 + 4 * ks + lk, h);
    };
    double yacc[4][G];
#pragma unroll
    for (int rg = 0; rg < 4; ++rg)
#pragma unroll
        for (int h = 0; h < G; ++h) yacc[rg][h] = 0.0;
    fetch(Ja);
    for (int J = Ja; J < Jb; ++J) {
        double hc[16], xj[4][G];
#pragma unroll
        for (int c = 0; c < 16; ++c) hc[c] = hn[c];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
#pragma unroll
            for (int h = 0; h < G; ++h) xj[ks][h] = xn[ks][h];
        if (J + 1 < Jb) fetch(J + 1);
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
#pragma unroll
            for (int rg = 0; rg < 4; ++rg)
#pragma unroll
                for (int h = 0; h < G; ++h)
                    yacc[rg][h] = __builtin_amdgcn_mfma_f64_4x4x4f64(hc[4 * rg + ks], xj[ks][h], yacc[rg][h], 0, 0, 0);
        if (J == I) continue;  // diagonal tile: the row sums cover it
        double zacc[4][G];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
#pragma unroll
            for (int h = 0; h < G; ++h) zacc[ks][h] = 0.0;
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
#pragma unroll
            for (int rg = 0; rg < 4; ++rg) {
                const double tr = __builtin_amdgcn_mfma_f64_4x4x4f64(hc[4 * rg + ks], ident, 0.0, 0, 0, 0);
#pragma unroll
                for (int h = 0; h < G; ++h)
                    zacc[ks][h] = __builtin_amdgcn_mfma_f64_4x4x4f64(tr, xi[rg][h], zacc[ks][h], 0, 0, 0);
            }
        // lane 16m + 4b + n, after the block sums: column 16w + 4ks + m, real
        // column 4h + n for every b; lane b stores ks = b
        const long tI = (long)I * (I - 1) / 2 + J;
        double* pz = part + 2 * (tI * TILE * R + (long)(16 * w + 4 * lb + lk) * R);
#pragma unroll
        for (int h = 0; h < G; ++h) {
            double v[4];
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) {
                const double z = zacc[ks][h] + dpp_f64<DPP_ROW_ROR8>(zacc[ks][h]);
                v[ks] = z + dpp_f64<DPP_ROW_ROR4>(z);
            }
            const double o = lb == 0 ? v[0] : lb == 1 ? v[1] : lb == 2 ? v[2] : v[3];
            if (4 * h + lq < C2) pz[4 * h + lq] = o;
        }
    }
    // row sums of wave w: lane 16m + 4b + n of (rg, h) is row 16rg + 4b + m, real column 4h + n
#pragma unroll
    for (int rg = 0; rg < 4; ++rg)
#pragma unroll
        for (int h = 0; h < G; ++h)
            if (4 * h + lq < C2) ys[(w * TILE + 16 * rg + 4 * lb + lk) * C2 + 4 * h + lq] = yacc[rg][h];
    __syncthreads();
    const long pcol = (long)nb * (nb - 1) / 2 * TILE * R;
    double* py = part + 2 * (pcol + (long)blockIdx.x * TILE * R);
    for (int e = t; e < TILE * C2; e += 256)
        py[e] = ((ys[e] + ys[TILE * C2 + e]) + ys[2 * TILE * C2 + e]) + ys[3 * TILE * C2 + e];
}

template <int R>
static void launch_symm3(const MatvecArgs& a, hipStream_t st) {
    const int nb = a.ld / TILE;
    const int S = sym_seg_for(nb, 3);
    hipLaunchKernelGGL((symm3_tiles_kernel<R>), dim3(sym_segment_count(nb, S)), dim3(256), 0, st, a.M,
                       reinterpret_cast<const double*>(a.x), reinterpret_cast<double*>(a.partial), a.ld, a.N,
                       a.x_stride, a.x_off, S);
    launch_sym_reduce<R>(a, nb, S, st);
}

// ------------------------------------------------------------------ symmetric, 2 x 2 super-tiles
// Variant 19: the 4x4x4 matrix-core tile products of variant 18, with a
// workgroup owning the super-tile (P, Q), Q <= P: block rows {2P, 2P+1} x
// block columns {2Q, 2Q+1} of the lower triangle, up to four tiles, processed
// one after the other with the next tile's loads in flight.  Row sums collect
// over the two tiles of each block row and column sums over the two tiles of
// each block column before they are written, so there are (up to) 2 + 2
// partials per four tiles, against 1 + 1/S per tile for row segments.
// Partials: Pcol[P^2 + J] (64 x R, J <= 2P), then Prow[2 (P (P + 1) / 2 + Q) + a].
long sym2_partial_cplx(int ld, int R) {
    const long nb = ld / TILE, np = (nb + 1) / 2;
    return (np * np + np * (np + 1)) * TILE * R;
}

template <int R>
__global__ __launch_bounds__(256) void symm2_tiles_kernel(const double* __restrict__ Mg, const double* __restrict__ xg,
                                                          double* __restrict__ part, int ld, int N, int x_stride,
                                                          int x_off) {
    constexpr int HP = TILE + 1;
    constexpr int C2 = 2 * R;
    constexpr int XE = TILE * C2;
    constexpr int XPT = (XE + 255) / 256;
    constexpr int G = (C2 + 3) / 4;
    __shared__ double Hs[TILE * HP];
    __shared__ double xs[2 * XE];   // x of block columns 2Q, 2Q + 1
    __shared__ double xis[2 * XE];  // x of block rows 2P, 2P + 1
    const int nb = ld / TILE, npr = (nb + 1) / 2;
    const int idx = blockIdx.x;
    int P = (int)((sqrt(8.0 * idx + 1.0) - 1.0) * 0.5);
    while (P * (P + 1) / 2 > idx) --P;
    while ((P + 1) * (P + 2) / 2 <= idx) ++P;
    const int Q = idx - P * (P + 1) / 2;
    const int t = threadIdx.x, lane = t & 63;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int lr = lane & 15, lk = lane >> 4, xq = lane & 3;
    const long xs2 = 2L * x_stride, xo2 = 2L * x_off;
    const bool hasB = 2 * P + 1 < nb;
    // valid tiles k = 2a + b: (2P + a, 2Q + b) inside the matrix and on or below the diagonal
    const unsigned valid = 1u | ((Q < P ? 1u : 0u) << 1) | ((hasB ? 1u : 0u) << 2) | ((hasB ? 1u : 0u) << 3);
    // x blocks (a missing block is clamped to the last one and never used); H's
    // padding is [0 0; 0 I] and x rows >= N are read as row N - 1
#pragma unroll
    for (int blk = 0; blk < 2; ++blk) {
        const int rI = min(2 * P + blk, nb - 1) * TILE, rJ = min(2 * Q + blk, nb - 1) * TILE;
#pragma unroll
        for (int u = 0; u < XPT; ++u) {
            const int e = t + 256 * u;
            if (XE % 256 == 0 || e < XE) {
                xis[blk * XE + e] = xg[(long)min(rI + e / C2, N - 1) * xs2 + xo2 + e % C2];
                xs[blk * XE + e] = xg[(long)min(rJ + e / C2, N - 1) * xs2 + xo2 + e % C2];
            }
        }
    }
    double hn[16];
    auto fetch = [&](int k) {
        const int i0 = (2 * P + (k >> 1)) * TILE, j0 = (2 * Q + (k & 1)) * TILE;
#pragma unroll
        for (int c = 0; c < 16; ++c) hn[c] = Mg[(long)(j0 + 4 * c + lk) * ld + i0 + 16 * w + lr];
    };
    double yacc[2][G], zacc[2][G];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int h = 0; h < G; ++h) yacc[a][h] = zacc[a][h] = 0.0;
    fetch(0);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (!((valid >> k) & 1u)) continue;  // uniform over the workgroup
        const int a = k >> 1, b = k & 1;
        const bool diag = 2 * P + a == 2 * Q + b;
        double hc[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) hc[c] = hn[c];
        __syncthreads();  // previous tile's LDS readers are done (and x staged, first time)
#pragma unroll
        for (int c = 0; c < 16; ++c) Hs[(4 * c + lk) * HP + 16 * w + lr] = hc[c];
        int kn = k + 1;
        while (kn < 4 && !((valid >> kn) & 1u)) ++kn;
        if (kn < 4) fetch(kn);
        __syncthreads();  // tile staged
#pragma unroll
        for (int c = 0; c < 16; ++c)
#pragma unroll
            for (int h = 0; h < G; ++h)
                yacc[a][h] = __builtin_amdgcn_mfma_f64_4x4x4f64(
                    hc[c], xs[b * XE + (4 * c + lk) * C2 + min(4 * h + xq, C2 - 1)], yacc[a][h], 0, 0, 0);
        if (diag) continue;  // diagonal tile: the row sums cover it
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const double hv = Hs[(16 * w + lr) * HP + 4 * r + lk];
#pragma unroll
            for (int h = 0; h < G; ++h)
                zacc[b][h] = __builtin_amdgcn_mfma_f64_4x4x4f64(
                    hv, xis[a * XE + (4 * r + lk) * C2 + min(4 * h + xq, C2 - 1)], zacc[b][h], 0, 0, 0);
        }
    }
    // D lane l: row (column) 16w + 4 ((l >> 2) & 3) + (l >> 4) of its block, real column 4h + (l & 3)
    const int rowl = 16 * w + 4 * ((lane >> 2) & 3) + lk;
    const long E = (long)TILE * R;
#pragma unroll
    for (int b = 0; b < 2; ++b) {
        if (!(Q < P || (b == 0 && hasB))) continue;  // no off-diagonal tile in column 2Q + b
        double* pz = part + 2 * (((long)P * P + 2 * Q + b) * E + (long)rowl * R);
#pragma unroll
        for (int h = 0; h < G; ++h)
            if (4 * h + xq < C2) pz[4 * h + xq] = zacc[b][h];
    }
    const long prow = (long)npr * npr * E;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        if (a == 1 && !hasB) continue;
        double* py = part + 2 * (prow + ((long)idx * 2 + a) * E + (long)rowl * R);
#pragma unroll
        for (int h = 0; h < G; ++h)
            if (4 * h + xq < C2) py[4 * h + xq] = yacc[a][h];
    }
}

// Phase 2 of variant 19: block B adds the row partials of super-tiles
// (B / 2, 0..B / 2) (row B % 2 of each) and the column partials Pcol[P^2 + B]
// for the super-tile rows P below B, in a fixed order, then the epilogue.
template <int R>
__global__ __launch_bounds__(SYM_REDUCE_THREADS) void sym2_reduce_kernel(MatvecArgs a) {
    __shared__ cplx red[SYM_G * 64];
    const int nb = a.ld / TILE, npr = (nb + 1) / 2;
    const int B = blockIdx.x, chunk = blockIdx.y;
    const int t = threadIdx.x, g = t >> 6, e = chunk * 64 + (t & 63);
    const long E = (long)TILE * R;
    const long prow = (long)npr * npr * E;
    const int P = B / 2, arow = B % 2;
    const int plo = (B % 2 == 0 && B + 1 < nb) ? B / 2 : B / 2 + 1;  // first super-tile row with a column partial
    const int nr = P + 1;
    const int L = nr + (npr - plo);
    const int s0 = g * L / SYM_G, s1 = (g + 1) * L / SYM_G;
    const int i = B * TILE + e / R;
    const long oi = (long)i * a.out_stride + a.out_off + e % R;
    EpiOperands ep{};
    if (t < 64 && i < a.N) ep = epilogue_operands(a, i, e % R);
    double sr = 0.0, si = 0.0;
    for (int base = s0; base < s1; base += SYM_MAXS) {
        cplx p[SYM_MAXS];
#pragma unroll
        for (int u = 0; u < SYM_MAXS; ++u) {
            const int src = min(base + u, s1 - 1);
            const long P2 = plo + (src - nr);
            const long off = src < nr ? prow + ((long)(P * (P + 1) / 2 + src) * 2 + arow) * E : (P2 * P2 + B) * E;
            p[u] = a.partial[off + e];
        }
#pragma unroll
        for (int u = 0; u < SYM_MAXS; ++u)
            if (base + u < s1) {
                sr += p[u].re;
                si += p[u].im;
            }
    }
    red[t] = cplx{sr, si};
    __syncthreads();
    if (t < 64) {
        cplx y{0.0, 0.0};
#pragma unroll
        for (int gg = 0; gg < SYM_G; ++gg) {
            y.re += red[gg * 64 + t].re;
            y.im += red[gg * 64 + t].im;
        }
        if (i < a.N) term_epilogue_pre(a, i, e % R, y, ep);
    }
}

template <int R>
static void launch_symm2(const MatvecArgs& a, hipStream_t st) {
    const int nb = a.ld / TILE, npr = (nb + 1) / 2;
    hipLaunchKernelGGL((symm2_tiles_kernel<R>), dim3(npr * (npr + 1) / 2), dim3(256), 0, st, a.M,
                       reinterpret_cast<const double*>(a.x), reinterpret_cast<double*>(a.partial), a.ld, a.N,
                       a.x_stride, a.x_off);
    hipLaunchKernelGGL((sym2_reduce_kernel<R>), dim3(nb, R), dim3(SYM_REDUCE_THREADS), 0, st, a);
}

// ------------------------------------------------------------------ symmetric, loads up front
// Variant 21: the segments, partial layout and 4x4x4 matrix-core tile
// products of variant 18 (so sym_reduce_kernel is its phase 2), with a fixed
// segment length T = 3 and every tile and x_J block of the segment loaded
// into registers before the first product: a workgroup has its whole segment
// (3 x 32 KB of H) in flight at once instead of one tile ahead, and the grid
// may exceed the resident workgroups (short-lived workgroups replace each
// other).  At N = 4096 this is within 2 % of variant 18 (segments of 2 or 4
// were slower): the term is not bound by the bytes in flight.
template <int R, int T>
__global__ __launch_bounds__(256) void symm_up_kernel(const double* __restrict__ Mg, const double* __restrict__ xg,
                                                      double* __restrict__ part, int ld, int N, int x_stride,
                                                      int x_off) {
    constexpr int HP = TILE + 1;
    constexpr int C2 = 2 * R;
    constexpr int XE = TILE * C2;
    constexpr int XPT = (XE + 255) / 256;
    constexpr int G = (C2 + 3) / 4;
    __shared__ double Hs[TILE * HP];
    __shared__ double xs[XE];
    __shared__ double xis[XE];
    const int nb = ld / TILE;
    const int seg = blockIdx.x;
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sym_segment_start(mid, T) <= seg) lo = mid;
        else hi = mid - 1;
    }
    const int I = lo;
    const int Ja = (seg - sym_segment_start(I, T)) * T, Jb = min(I + 1, Ja + T);
    const int t = threadIdx.x, lane = t & 63;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int i0 = I * TILE;
    const int lr = lane & 15, lk = lane >> 4, xq = lane & 3;
    const long xs2 = 2L * x_stride, xo2 = 2L * x_off;
    // H's padding is [0 0; 0 I] and x rows >= N are read as row N - 1 (finite
    // data): rows >= N of x never reach rows < N of the result.
    double hv[T][16], xv[T][XPT], xi[XPT];
#pragma unroll
    for (int u = 0; u < T; ++u) {
        if (Ja + u < Jb) {
            const int j0 = (Ja + u) * TILE;
#pragma unroll
            for (int c = 0; c < 16; ++c) hv[u][c] = Mg[(long)(j0 + 4 * c + lk) * ld + i0 + 16 * w + lr];
#pragma unroll
            for (int q = 0; q < XPT; ++q) {
                const int e = t + 256 * q;
                if (XE % 256 == 0 || e < XE) xv[u][q] = xg[(long)min(j0 + e / C2, N - 1) * xs2 + xo2 + e % C2];
            }
        }
    }
#pragma unroll
    for (int q = 0; q < XPT; ++q) {
        const int e = t + 256 * q;
        if (XE % 256 == 0 || e < XE) xi[q] = xg[(long)min(i0 + e / C2, N - 1) * xs2 + xo2 + e % C2];
    }
#pragma unroll
    for (int q = 0; q < XPT; ++q) {
        const int e = t + 256 * q;
        if (XE % 256 == 0 || e < XE) xis[e] = xi[q];
    }
    double yacc[G];
#pragma unroll
    for (int h = 0; h < G; ++h) yacc[h] = 0.0;
    // D lane l: row (column) 16w + 4 ((l >> 2) & 3) + (l >> 4) of the tile, real column 4h + (l & 3)
    const int rowl = 16 * w + 4 * ((lane >> 2) & 3) + lk;
#pragma unroll
    for (int u = 0; u < T; ++u) {
        const int J = Ja + u;
        if (J >= Jb) break;  // uniform over the workgroup
        __syncthreads();     // previous tile's LDS readers are done (x_I staged, first time)
#pragma unroll
        for (int c = 0; c < 16; ++c) Hs[(4 * c + lk) * HP + 16 * w + lr] = hv[u][c];
#pragma unroll
        for (int q = 0; q < XPT; ++q) {
            const int e = t + 256 * q;
            if (XE % 256 == 0 || e < XE) xs[e] = xv[u][q];
        }
        __syncthreads();     // tile and x_J staged
#pragma unroll
        for (int c = 0; c < 16; ++c)
#pragma unroll
            for (int h = 0; h < G; ++h)
                yacc[h] = __builtin_amdgcn_mfma_f64_4x4x4f64(hv[u][c], xs[(4 * c + lk) * C2 + min(4 * h + xq, C2 - 1)],
                                                             yacc[h], 0, 0, 0);
        if (J == I) continue;  // diagonal tile: the row sums cover it
        double zacc[G];
#pragma unroll
        for (int h = 0; h < G; ++h) zacc[h] = 0.0;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const double hz = Hs[(16 * w + lr) * HP + 4 * r + lk];
#pragma unroll
            for (int h = 0; h < G; ++h)
                zacc[h] = __builtin_amdgcn_mfma_f64_4x4x4f64(hz, xis[(4 * r + lk) * C2 + min(4 * h + xq, C2 - 1)],
                                                             zacc[h], 0, 0, 0);
        }
        const long tI = (long)I * (I - 1) / 2 + J;
        double* pz = part + 2 * (tI * TILE * R + (long)rowl * R);
#pragma unroll
        for (int h = 0; h < G; ++h)
            if (4 * h + xq < C2) pz[4 * h + xq] = zacc[h];
    }
    const long pcol = (long)nb * (nb - 1) / 2 * TILE * R;
    double* py = part + 2 * (pcol + (long)seg * TILE * R + (long)rowl * R);
#pragma unroll
    for (int h = 0; h < G; ++h)
        if (4 * h + xq < C2) py[4 * h + xq] = yacc[h];
}

template <int R, int T>
static void launch_symm_up(const MatvecArgs& a, hipStream_t st) {
    const int nb = a.ld / TILE;
    hipLaunchKernelGGL((symm_up_kernel<R, T>), dim3(sym_segment_count(nb, T)), dim3(256), 0, st, a.M,
                       reinterpret_cast<const double*>(a.x), reinterpret_cast<double*>(a.partial), a.ld, a.N,
                       a.x_stride, a.x_off);
    launch_sym_reduce<R>(a, nb, T, st);
}

constexpr int SYM_UP_T = 3;  // segment length of variant 21

static void launch_symm_up_variant(const MatvecArgs& a, hipStream_t st) {
    switch (a.R) {
        case 1: launch_symm_up<1, SYM_UP_T>(a, st); break;
        case 2: launch_symm_up<2, SYM_UP_T>(a, st); break;
        case 4: launch_symm_up<4, SYM_UP_T>(a, st); break;
        default: launch_symm_up<8, SYM_UP_T>(a, st); break;
    }
}

// Variant table: 0 = split-K with last-block reduction (needs a.split, partial,
// counters); 1..4 and 9..14 = column-block kernels of different shapes and LDS
// schedules; 5..8 = two-phase split-K; 15 = split-K matrix-core GEMM (full
// operator); 16 = symmetric half storage, vector FMAs; 17 = symmetric half
// storage, matrix-core tile products (16x16x4); 18 = the same with 4x4x4 products;
// 19 = 4x4x4 products over 2 x 2 super-tiles; 20 = 4x4x4 products over
// per-wave column strips, without LDS staging of the tiles; 21 = the
// products of 18 over segments of 3 tiles, all loaded up front.
constexpr int MV_VARIANTS = 22;

template <int R>
static void launch_mv_r(const MatvecArgs& a, int variant, hipStream_t st) {
    // LDS tile height: a multiple of 64 rows, 64 KB of double-buffered x at most
    constexpr int JC = ((2048 / R) > 512 ? 512 : (2048 / R)) & ~63;
    constexpr int CS = R <= 4 ? 4 : (R <= 8 ? 2 : 1);
    switch (variant) {
        case 0: launch_mv_split<R, CS>(a, st); break;
        case 1: launch_mv_cols<R, 2, 8, JC>(a, st); break;
        case 2: launch_mv_cols<R, 4, 4, JC>(a, st); break;
        case 3: launch_mv_cols<R, 1, 8, JC>(a, st); break;
        case 4: launch_mv_cols<R, 2, 4, JC>(a, st); break;
        // two-phase split-K; the split keeps split * N * R within the partial buffer
        case 5: launch_mv_twophase<R, CS, 2>(a, a.split, st); break;
        case 6: launch_mv_twophase<R, CS, 4>(a, a.split, st); break;
        case 7: launch_mv_twophase<R, CS, 2>(a, a.split > 1 ? (a.split + 1) / 2 : 1, st); break;
        case 8: launch_mv_twophase<R, (CS > 1 ? CS / 2 : 1), 2>(a, a.split, st); break;
        // column-block kernels with a row-by-row LDS schedule (lower register use)
        case 9: launch_mv_cols<R, 2, 8, JC, 1>(a, st); break;
        case 10: launch_mv_cols<R, 2, 8, JC, 2>(a, st); break;
        case 11: launch_mv_cols<R, 1, 16, JC, 1>(a, st); break;
        case 12: launch_mv_cols<R, 2, 8, (JC >= 512 ? 256 : JC), 1>(a, st); break;
        case 13: launch_mv_cols<R, 4, 4, JC, 1>(a, st); break;
        default: launch_mv_cols<R, 1, 8, JC, 1>(a, st); break;
    }
}

static int g_variant_override = -1;
struct TuneKey {
    int N, R, tri;
};
// Process-wide: a shape is tuned once, by the first system that has it, so
// every later system of that shape runs the same kernels (same summation order).
static std::vector<std::pair<TuneKey, int>> g_tuned;
static std::mutex g_tuned_mu;

int matvec_variant_for(int N, int R, int tri) {
    if (g_variant_override >= 0) return g_variant_override;
    std::lock_guard<std::mutex> lk(g_tuned_mu);
    for (auto& kv : g_tuned)
        if (kv.first.N == N && kv.first.R == R && kv.first.tri == tri) return kv.second;
    return 1;
}
bool matvec_is_tuned(int N, int R, int tri) {
    std::lock_guard<std::mutex> lk(g_tuned_mu);
    for (auto& kv : g_tuned)
        if (kv.first.N == N && kv.first.R == R && kv.first.tri == tri) return true;
    return false;
}
void matvec_set_tuned(int N, int R, int tri, int variant) {
    std::lock_guard<std::mutex> lk(g_tuned_mu);
    for (auto& kv : g_tuned)
        if (kv.first.N == N && kv.first.R == R && kv.first.tri == tri) {
            kv.second = variant;
            return;
        }
    g_tuned.push_back({TuneKey{N, R, tri}, variant});
}
void matvec_force_variant(int v) { g_variant_override = v; }
int matvec_variant_count() { return MV_VARIANTS; }

void launch_matvec_variant(const MatvecArgs& a, int variant, hipStream_t st) {
    if (variant == 21) {
        if (!sym_supported(a.R, a.tri)) {
            launch_matvec_variant(a, 1, st);
            return;
        }
        launch_symm_up_variant(a, st);
        return;
    }
    if (variant == 19) {
        if (!sym_supported(a.R, a.tri)) {
            launch_matvec_variant(a, 1, st);
            return;
        }
        switch (a.R) {
            case 1: launch_symm2<1>(a, st); break;
            case 2: launch_symm2<2>(a, st); break;
            case 4: launch_symm2<4>(a, st); break;
            default: launch_symm2<8>(a, st); break;
        }
        return;
    }
    if (variant == 20) {
        if (!sym_supported(a.R, a.tri)) {
            launch_matvec_variant(a, 1, st);
            return;
        }
        switch (a.R) {
            case 1: launch_symm3<1>(a, st); break;
            case 2: launch_symm3<2>(a, st); break;
            case 4: launch_symm3<4>(a, st); break;
            default: launch_symm3<8>(a, st); break;
        }
        return;
    }
    if (variant == 17 || variant == 18) {
        // symmetric kernel with matrix-core tile products (16x16x4 or 4x4x4);
        // other shapes use variant 1
        if (!sym_supported(a.R, a.tri)) {
            launch_matvec_variant(a, 1, st);
            return;
        }
        const bool small = variant == 18;
        switch (a.R) {
            case 1: small ? launch_symm<1, 1>(a, st) : launch_symm<1, 0>(a, st); break;
            case 2: small ? launch_symm<2, 1>(a, st) : launch_symm<2, 0>(a, st); break;
            case 4: small ? launch_symm<4, 1>(a, st) : launch_symm<4, 0>(a, st); break;
            default: small ? launch_symm<8, 1>(a, st) : launch_symm<8, 0>(a, st); break;
        }
        return;
    }
    if (variant == 16) {
        // symmetric half-storage two-phase kernel; other shapes use the column-block kernel
        if (!sym_supported(a.R, a.tri)) {
            launch_matvec_variant(a, 1, st);
            return;
        }
        switch (a.R) {
            case 1: launch_sym<1>(a, st); break;
            case 2: launch_sym<2>(a, st); break;
            case 4: launch_sym<4>(a, st); break;
            default: launch_sym<8>(a, st); break;
        }
        return;
    }
    if (variant == 15) {
        // matrix-core split-K GEMM; triangular shapes use the column-block kernel
        if (a.tri == TRI_FULL) launch_gemm3(a, 0, reinterpret_cast<double*>(a.partial), st);
        else launch_matvec_variant(a, 9, st);
        return;
    }
    switch (a.R) {
        case 1: launch_mv_r<1>(a, variant, st); break;
        case 2: launch_mv_r<2>(a, variant, st); break;
        case 3: launch_mv_r<3>(a, variant, st); break;
        case 4: launch_mv_r<4>(a, variant, st); break;
        case 6: launch_mv_r<6>(a, variant, st); break;
        case 8: launch_mv_r<8>(a, variant, st); break;
        case 12: launch_mv_r<12>(a, variant, st); break;
        case 16: launch_mv_r<16>(a, variant, st); break;
        default: {
            int done = 0;
            while (done < a.R) {
                const int take = a.R - done;
                const int t = take >= 16 ? 16 : take >= 12 ? 12 : take >= 8 ? 8 : take >= 6 ? 6 : take >= 4 ? 4 : take >= 3 ? 3 : take >= 2 ? 2 : 1;
                MatvecArgs b = a;
                b.R = t;
                b.x_off = a.x_off + done;
                b.out_off = a.out_off + done;
                launch_matvec_variant(b, variant, st);
                done += t;
            }
        }
    }
}

void launch_matvec(const MatvecArgs& a, hipStream_t st) {
    launch_matvec_variant(a, matvec_variant_for(a.N, a.R, a.tri), st);
}

// Term on the FP32-stored operator a.M32 with FP64 arithmetic: the symmetric
// half-storage kernel with 4x4x4 matrix-core products for R <= 8, else the
// split-K GEMM with FP64 products.  One fixed kernel per shape, so results are
// reproducible run to run.
void launch_term_f32h(const MatvecArgs& a, hipStream_t st) {
    if (!sym_supported(a.R, a.tri)) {
        launch_gemm3(a, 2, reinterpret_cast<double*>(a.partial), st);
        return;
    }
    switch (a.R) {
        case 1: launch_symm<1, 1, float>(a, st); break;
        case 2: launch_symm<2, 1, float>(a, st); break;
        case 4: launch_symm<4, 1, float>(a, st); break;
        default: launch_symm<8, 1, float>(a, st); break;
    }
}

// ------------------------------------------------------------------ Lanczos
// One Lanczos step on a single workgroup (1024 threads):
//   alpha_j = v.w ; w -= alpha_j v + beta_j v_prev ; beta_{j+1} = |w| ;
//   v_prev = v ; v = w / beta_{j+1}.
// w is the real part of a complex buffer (stride 1 complex).
__device__ double block_sum_1024(double v, double* sh) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x < 64) {
        t = threadIdx.x < (int)(blockDim.x >> 6) ? sh[threadIdx.x] : 0.0;
        t = wave_sum(t);
    }
    __syncthreads();
    if (threadIdx.x == 0) sh[0] = t;
    __syncthreads();
    const double r = sh[0];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(1024) void lanczos_step_kernel(int N, cplx* v, cplx* vprev, const cplx* w, double* alpha,
                                                            double* beta, int j) {
    __shared__ double sh[64];
    double p = 0.0;
    for (int i = threadIdx.x; i < N; i += blockDim.x) p += v[i].re * w[i].re;
    const double al = block_sum_1024(p, sh);
    const double be = j > 0 ? beta[j] : 0.0;
    double q = 0.0;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        const double wi = w[i].re - al * v[i].re - be * vprev[i].re;
        q += wi * wi;
    }
    const double nb = sqrt(block_sum_1024(q, sh));
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        const double wi = w[i].re - al * v[i].re - be * vprev[i].re;
        const double vi = v[i].re;
        vprev[i] = cplx{vi, 0.0};
        v[i] = cplx{nb > 0.0 ? wi / nb : 0.0, 0.0};
    }
    if (threadIdx.x == 0) {
        alpha[j] = al;
        beta[j + 1] = nb;
    }
}

__global__ void lanczos_start_kernel(int N, cplx* v, cplx* vprev, unsigned seed) {
    // deterministic pseudo-random start vector, normalised below by the first step's norm
    __shared__ double sh[64];
    double q = 0.0;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        unsigned x = (unsigned)i * 2654435761u ^ seed;
        x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
        const double r = ((double)x / 4294967296.0) - 0.5;
        v[i] = cplx{r, 0.0};
        vprev[i] = cplx{0.0, 0.0};
        q += r * r;
    }
    const double nrm = sqrt(block_sum_1024(q, sh));
    for (int i = threadIdx.x; i < N; i += blockDim.x) v[i].re /= nrm;
}

// Sturm count: number of eigenvalues of the m x m tridiagonal below x.
static int sturm_count(const double* al, const double* be, int m, double x) {
    int cnt = 0;
    double q = 1.0;
    for (int k = 0; k < m; ++k) {
        const double b2 = k > 0 ? be[k] * be[k] : 0.0;
        q = (al[k] - x) - (k > 0 ? b2 / q : 0.0);
        if (q == 0.0) q = -1e-300;
        if (q < 0.0) ++cnt;
    }
    return cnt;
}

// Extreme Ritz values by bisection, residual of each extreme Ritz pair
// (|beta_m s_m|, s from inverse iteration on the tridiagonal), and the
// enclosing interval [emin, emax] = [theta_min - margin_lo, theta_max + margin_hi].
// Runs on the host: the m <= 120 recurrence coefficients come back with one
// copy, and the serial bisection takes microseconds there against ~0.6 ms on
// one GPU thread.
void lanczos_bounds_host(const double* al, const double* be, int m, double rel_margin, double* bounds) {
    double lo = 1e300, hi = -1e300;
    for (int k = 0; k < m; ++k) {
        const double r = (k > 0 ? std::fabs(be[k]) : 0.0) + (k + 1 < m ? std::fabs(be[k + 1]) : 0.0);
        lo = std::fmin(lo, al[k] - r);
        hi = std::fmax(hi, al[k] + r);
    }
    double th[2];
    for (int e = 0; e < 2; ++e) {
        const int target = e == 0 ? 1 : m;  // smallest: count(x) >= 1 ; largest: count(x) >= m
        double a = lo, b = hi;
        for (int it = 0; it < 200; ++it) {
            const double mid = 0.5 * (a + b);
            if (sturm_count(al, be, m, mid) >= target) b = mid; else a = mid;
            if (b - a <= 1e-13 * std::fmax(1.0, std::fabs(mid))) break;
        }
        th[e] = 0.5 * (a + b);
    }
    double res[2];
    for (int e = 0; e < 2; ++e) {
        // inverse iteration with a slightly shifted theta on the tridiagonal
        const double shift = th[e] + (e == 0 ? -1e-10 : 1e-10) * std::fmax(1.0, std::fabs(th[e]));
        double s[128], dd[128], uu[128];
        for (int k = 0; k < m; ++k) s[k] = 1.0;
        for (int it = 0; it < 3; ++it) {
            // LU of (T - shift I) without pivoting (Thomas algorithm)
            dd[0] = al[0] - shift;
            for (int k = 1; k < m; ++k) {
                const double l = be[k] / (std::fabs(dd[k - 1]) > 1e-300 ? dd[k - 1] : 1e-300);
                dd[k] = al[k] - shift - l * be[k];
                uu[k] = l;
            }
            for (int k = 1; k < m; ++k) s[k] -= uu[k] * s[k - 1];
            s[m - 1] /= (std::fabs(dd[m - 1]) > 1e-300 ? dd[m - 1] : 1e-300);
            for (int k = m - 2; k >= 0; --k) s[k] = (s[k] - be[k + 1] * s[k + 1]) / (std::fabs(dd[k]) > 1e-300 ? dd[k] : 1e-300);
            double nrm = 0.0;
            for (int k = 0; k < m; ++k) nrm += s[k] * s[k];
            nrm = std::sqrt(nrm);
            for (int k = 0; k < m; ++k) s[k] /= nrm;
        }
        res[e] = std::fabs(be[m] * s[m - 1]);
    }
    const double width = th[1] - th[0];
    bounds[0] = th[0] - res[0] - rel_margin * width;
    bounds[1] = th[1] + res[1] + rel_margin * width;
    bounds[2] = th[0];
    bounds[3] = th[1];
    bounds[4] = res[0];
    bounds[5] = res[1];
}

// ------------------------------------------------------------------ Bessel coefficients
// J_k(x), k = 0..K, by Miller's backward recurrence J_{k-1} = (2k/x) J_k - J_{k+1}
// normalised with J_0 + 2 sum J_{2k} = 1; K = last k with |J_k| > tol.
// coef[k] = (2 - delta_k0) (-i)^k J_k(x) e^{-i c dt};  out_k[0] = K (number of terms - 1).
__global__ void bessel_coef_kernel(double x, double phase, int kmax_alloc, double tol, double* J, cplx* coef, int* out_k) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int K = 0;
    if (x < 1e-12) {
        J[0] = 1.0;
        K = 0;
    } else {
        int M = (int)(x + 15.0 * cbrt(x) + 40.0);
        M += M & 1;
        if (M > kmax_alloc - 2) M = kmax_alloc - 2;
        double jp1 = 0.0, jk = 1e-300;
        for (int k = 0; k <= M + 1; ++k) J[k] = 0.0;
        J[M] = jk;
        for (int k = M; k >= 1; --k) {
            const double jm1 = (2.0 * k / x) * jk - jp1;
            J[k - 1] = jm1;
            jp1 = jk;
            jk = jm1;
            if (fabs(jk) > 1e250) {
                for (int q = k - 1; q <= M; ++q) J[q] *= 1e-250;
                jk *= 1e-250;
                jp1 *= 1e-250;
            }
        }
        double norm = J[0];
        for (int k = 2; k <= M; k += 2) norm += 2.0 * J[k];
        for (int k = 0; k <= M; ++k) J[k] /= norm;
        K = M;
        while (K > 0 && fabs(J[K]) < tol) --K;
        K += 1;
        if (K > M) K = M;
    }
    const double cp = cos(phase), sp = -sin(phase);  // e^{-i phase}
    for (int k = 0; k <= K; ++k) {
        const double f = (k == 0 ? 1.0 : 2.0) * J[k];
        // (-i)^k
        double re, im;
        switch (k & 3) {
            case 0: re = f; im = 0.0; break;
            case 1: re = 0.0; im = -f; break;
            case 2: re = -f; im = 0.0; break;
            default: re = 0.0; im = f; break;
        }
        coef[k] = cplx{re * cp - im * sp, re * sp + im * cp};
    }
    coef[K + 1] = cplx{0.0, 0.0};
    out_k[0] = K;
}

// The same coefficients on the host (the recurrence above, operation for
// operation): coef[0..K] and coef[K + 1] = 0; returns K.
int bessel_coef_host(double x, double phase, int kmax_alloc, double tol, cplx* coef) {
    std::vector<double> J((size_t)kmax_alloc + 2, 0.0);
    int K = 0;
    if (x < 1e-12) {
        J[0] = 1.0;
    } else {
        int M = (int)(x + 15.0 * std::cbrt(x) + 40.0);
        M += M & 1;
        if (M > kmax_alloc - 2) M = kmax_alloc - 2;
        double jp1 = 0.0, jk = 1e-300;
        J[M] = jk;
        for (int k = M; k >= 1; --k) {
            const double jm1 = (2.0 * k / x) * jk - jp1;
            J[k - 1] = jm1;
            jp1 = jk;
            jk = jm1;
            if (std::fabs(jk) > 1e250) {
                for (int q = k - 1; q <= M; ++q) J[q] *= 1e-250;
                jk *= 1e-250;
                jp1 *= 1e-250;
            }
        }
        double norm = J[0];
        for (int k = 2; k <= M; k += 2) norm += 2.0 * J[k];
        for (int k = 0; k <= M; ++k) J[k] /= norm;
        K = M;
        while (K > 0 && std::fabs(J[K]) < tol) --K;
        K += 1;
        if (K > M) K = M;
    }
    const double cp = std::cos(phase), sp = -std::sin(phase);
    for (int k = 0; k <= K; ++k) {
        const double f = (k == 0 ? 1.0 : 2.0) * J[k];
        double re, im;
        switch (k & 3) {
            case 0: re = f; im = 0.0; break;
            case 1: re = 0.0; im = -f; break;
            case 2: re = -f; im = 0.0; break;
            default: re = 0.0; im = f; break;
        }
        coef[k] = cplx{re * cp - im * sp, re * sp + im * cp};
    }
    coef[K + 1] = cplx{0.0, 0.0};
    return K;
}

// ------------------------------------------------------------------ observables
// One workgroup per wavepacket w: fragment populations
//   pop_f = Re sum_{i in f} conj(ket_i) bra_i  (data_output.f:242-263, DUAL_bra = conj(ket), DUAL_ket = bra)
// and <H> = sum_i conj(phi_bra_i) (H_orth phi_ket)_i (ElHl_Chebyshev.f:329-371).
// out: [bra (W) | AO_bra (W) | ket (W)] per row; phi: [bra (W) | ket (W)]; hk: N x W.
// The fragment sums stay in registers: a row adds its population to every
// fragment's sum, zero to all but its own (x + 0 = x), so no array is indexed
// by the row's fragment (a dynamically indexed array would live in scratch
// memory).  The loop is unrolled so that several rows' loads are in flight;
// each thread's sums keep their row order.
__device__ __forceinline__ void observables_body(int N, int W, int n_frag, const int* frag, const cplx* out,
                                                 const cplx* phi, const cplx* hk, double* pops, double* erg,
                                                 double* norms, const int w) {
    constexpr int MAXF = 16;
    __shared__ double sh[(MAXF + 5) * 4];
    double pf[MAXF + 1];
#pragma unroll
    for (int f = 0; f <= MAXF; ++f) pf[f] = 0.0;
    double er = 0.0, ei = 0.0, nk = 0.0, nb = 0.0;
#pragma unroll 4
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        const cplx b = out[(long)i * 3 * W + w];
        const cplx k = out[(long)i * 3 * W + 2 * W + w];
        const double p = k.re * b.re + k.im * b.im;  // Re(conj(k) b)
        const int f = frag[i];
#pragma unroll
        for (int ff = 0; ff < MAXF; ++ff) pf[ff] += f == ff ? p : 0.0;
        pf[MAXF] += p;
        const cplx pb = phi[(long)i * 2 * W + w];
        const cplx pk = phi[(long)i * 2 * W + W + w];
        const cplx hv = hk[(long)i * W + w];
        er += pb.re * hv.re + pb.im * hv.im;
        ei += pb.re * hv.im - pb.im * hv.re;
        nb += pb.re * pb.re + pb.im * pb.im;
        nk += pk.re * pk.re + pk.im * pk.im;
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    auto red = [&](double v, int slot) {
        v = wave_sum(v);
        if (lane == 0) sh[slot * 4 + wv] = v;
    };
#pragma unroll
    for (int f = 0; f < MAXF; ++f)
        if (f < n_frag) red(pf[f], f);
    red(pf[MAXF], MAXF);
    red(er, MAXF + 1);
    red(ei, MAXF + 2);
    red(nb, MAXF + 3);
    red(nk, MAXF + 4);
    __syncthreads();
    if (threadIdx.x == 0) {
        auto tot = [&](int slot) { return ((sh[slot * 4 + 0] + sh[slot * 4 + 1]) + sh[slot * 4 + 2]) + sh[slot * 4 + 3]; };
        for (int f = 0; f < n_frag && f < MAXF; ++f) pops[w * (n_frag + 1) + f] = tot(f);
        pops[w * (n_frag + 1) + n_frag] = tot(MAXF);
        erg[2 * w] = tot(MAXF + 1);
        erg[2 * w + 1] = tot(MAXF + 2);
        norms[2 * w] = tot(MAXF + 3);
        norms[2 * w + 1] = tot(MAXF + 4);
    }
}

__global__ __launch_bounds__(256) void observables_kernel(int N, int W, int n_frag, const int* frag, const cplx* out,
                                                          const cplx* phi, const cplx* hk, double* pops, double* erg,
                                                          double* norms) {
    observables_body(N, W, n_frag, frag, out, phi, hk, pops, erg, norms, blockIdx.x);
}

__global__ __launch_bounds__(256) void grouped_observables_kernel(const ObsTask* tasks) {
    const ObsTask t = tasks[blockIdx.x];
    observables_body(t.N, t.W, t.n_frag, t.frag, t.out, t.phi, t.hk, t.pops, t.erg, t.norms, t.w);
}

// Column norms |phi_c|^2 of an N x R complex block (one workgroup per column).
__global__ __launch_bounds__(256) void column_norms_kernel(int N, int R, const cplx* x, double* out) {
    __shared__ double sh[4];
    const int c = blockIdx.x;
    double s = 0.0;
#pragma unroll 4
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        const cplx v = x[(long)i * R + c];
        s += v.re * v.re + v.im * v.im;
    }
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[c] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

__global__ __launch_bounds__(256) void grouped_norms_kernel(const NormTask* tasks) {
    __shared__ double sh[4];
    const NormTask t = tasks[blockIdx.x];
    double s = 0.0;
#pragma unroll 4
    for (int i = threadIdx.x; i < t.N; i += blockDim.x) {
        const cplx v = t.x[(long)i * t.R + t.c];
        s += v.re * v.re + v.im * v.im;
    }
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) *t.out = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

__global__ void grouped_copy_kernel(const CopyTask* tasks) {
    const CopyTask t = tasks[blockIdx.x];
    for (long i = threadIdx.x; i < t.n; i += blockDim.x) t.dst[i] = t.src[i];
}

__global__ void zero_u32_kernel(unsigned* p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}
__global__ void zero_cplx_kernel(cplx* p, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = cplx{0.0, 0.0};
}


// phi[:, c] *= sqrt(n0[c] / n1[c]) (FP64 norm restoration after a mixed-precision step)
__global__ void renorm_columns_kernel(cplx* x, int N, int R, const double* n0, const double* n1) {
    const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (e >= (long)N * R) return;
    const int c = (int)(e % R);
    const double f = n1[c] > 0.0 ? sqrt(n0[c] / n1[c]) : 1.0;
    x[e].re *= f;
    x[e].im *= f;
}

// H32 = fl32(H) and the residual H - H32 in bf16, both column-major like H
__global__ void split_f32_bf16_kernel(const double* __restrict__ H, float* __restrict__ H32,
                                      unsigned short* __restrict__ Hlo, long n) {
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const double h = H[e];
        const float hi = (float)h;
        H32[e] = hi;
        Hlo[e] = bf16_rne((float)(h - (double)hi));
    }
}

__global__ void to_bf16_kernel(const float* __restrict__ a, unsigned short* __restrict__ b, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        b[i] = bf16_rne(a[i]);
}

__global__ void to_f32_kernel(const double* __restrict__ a, float* __restrict__ b, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}

// ------------------------------------------------------------------ launchers
void launch_grouped_norms(const NormTask* d_tasks, int n, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(grouped_norms_kernel, dim3(n), dim3(256), 0, st, d_tasks);
}
void launch_grouped_copy(const CopyTask* d_tasks, int n, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(grouped_copy_kernel, dim3(n), dim3(256), 0, st, d_tasks);
}
void launch_grouped_observables(const ObsTask* d_tasks, int n, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(grouped_observables_kernel, dim3(n), dim3(256), 0, st, d_tasks);
}
void launch_renorm_columns(cplx* x, int N, int R, const double* n0, const double* n1, hipStream_t st) {
    const long n = (long)N * R;
    hipLaunchKernelGGL(renorm_columns_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, N, R, n0, n1);
}
void launch_split_f32_bf16(const double* H, float* H32, unsigned short* Hlo, long n, hipStream_t st) {
    hipLaunchKernelGGL(split_f32_bf16_kernel, dim3(4096), dim3(256), 0, st, H, H32, Hlo, n);
}
void launch_to_bf16(const float* a, unsigned short* b, long n, hipStream_t st) {
    hipLaunchKernelGGL(to_bf16_kernel, dim3(1024), dim3(256), 0, st, a, b, n);
}
void launch_to_f32(const double* a, float* b, long n, hipStream_t st) {
    hipLaunchKernelGGL(to_f32_kernel, dim3(4096), dim3(256), 0, st, a, b, n);
}
void launch_zero_u32(unsigned* p, int n, hipStream_t st) {
    hipLaunchKernelGGL(zero_u32_kernel, dim3(64), dim3(256), 0, st, p, n);
}
void launch_zero_cplx(cplx* p, long n, hipStream_t st) {
    hipLaunchKernelGGL(zero_cplx_kernel, dim3(256), dim3(256), 0, st, p, n);
}
void launch_lanczos_start(int N, cplx* v, cplx* vprev, unsigned seed, hipStream_t st) {
    hipLaunchKernelGGL(lanczos_start_kernel, dim3(1), dim3(1024), 0, st, N, v, vprev, seed);
}
void launch_lanczos_step(int N, cplx* v, cplx* vprev, const cplx* w, double* alpha, double* beta, int j,
                         hipStream_t st) {
    hipLaunchKernelGGL(lanczos_step_kernel, dim3(1), dim3(1024), 0, st, N, v, vprev, w, alpha, beta, j);
}
void launch_bessel(double x, double phase, int kmax_alloc, double tol, double* J, cplx* coef, int* out_k,
                   hipStream_t st) {
    hipLaunchKernelGGL(bessel_coef_kernel, dim3(1), dim3(64), 0, st, x, phase, kmax_alloc, tol, J, coef, out_k);
}
void launch_observables(int N, int W, int n_frag, const int* frag, const cplx* out, const cplx* phi, const cplx* hk,
                        double* pops, double* erg, double* norms, hipStream_t st) {
    hipLaunchKernelGGL(observables_kernel, dim3(W), dim3(256), 0, st, N, W, n_frag, frag, out, phi, hk, pops, erg,
                       norms);
}
void launch_column_norms(int N, int R, const cplx* x, double* out, hipStream_t st) {
    hipLaunchKernelGGL(column_norms_kernel, dim3(R), dim3(256), 0, st, N, R, x, out);
}

}  // namespace dh
