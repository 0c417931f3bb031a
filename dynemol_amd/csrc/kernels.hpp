// Kernel launchers shared between the translation units of the library.
#pragma once
#include "common.hpp"

namespace dh {

constexpr int MAX_SHELLS = 3;  // s, p, d
constexpr int TILE = 64;       // dense-algebra block size; ld is padded to a multiple of it

// Device view of the basis (struct-of-arrays; all pointers are device memory).
struct BasisDev {
    int n_atoms;
    const double* coords;     // [n_atoms*3], x y z per atom, Angstrom
    const int* atom_nshell;   // [n_atoms]
    const int* atom_shell0;   // [n_atoms] first shell index
    const int* sh_n;          // [n_shells]
    const int* sh_l;
    const int* sh_nzeta;
    const int* sh_orb0;       // first orbital of the shell (m = -l)
    const double* sh_zeta;    // [2*n_shells], 1/Angstrom
    const double* sh_coef;    // [2*n_shells]
    const double* ip;         // [N] per orbital
    const double* kwh;        // [N]
    const double* vshift;     // [N]
    const int* atom_order;    // [n_atoms] atoms sorted by basis signature (groups equal elements)
};

// X_ij, i != j, of the extended-Hueckel Hamiltonian h_ij = X_ij S_ij
// (hamiltonians.f:33-63); X_ii = IP_i + V_shift_i
__host__ __device__ inline double huckel_x(double ipa, double ipb, double kwa, double kwb, double va, double vb) {
    const double c1 = ipa - ipb, c2 = ipa + ipb;
    const double c3 = (c1 / c2) * (c1 / c2);
    const double c4 = (va + vb) * 0.5;
    const double kwh = (kwa + kwb) * 0.5;
    const double keff = kwh + c3 + c3 * c3 * (1.0 - kwh);
    return keff * c2 * 0.5 + c4;
}

int launch_build_SH(const BasisDev& b, int N, int ld, double* S, double* h, double cutoff, double s_drop,
                    hipStream_t st);

// ---------------------------------------------------------------- dense FP64
enum KMode { K_FULL = 0, K_BELOW_ROWTILE = 1, K_BELOW_COLTILE = 2, K_FROM_ROWTILE = 3, K_FROM_COLTILE = 4 };

struct Gemm {
    int M, N, K;
    const double* A;
    long lda, strideA;
    const double* B;
    long ldb, strideB;
    double* C;
    long ldc, strideC;
    double alpha, beta;
    int transA, transB;
    int kmode;       // KMode: restrict k to where a triangular operand is nonzero
    int lower_only;  // compute only tiles with row-tile >= col-tile
    int batch;
    int pair;        // set by launch_gemm: 1 = pair row tiles i, n-1-i; 2 = column tiles
    int c_acc;       // set by launch_gemm for alpha = -1, beta = 1: C starts in the accumulators
};
void launch_gemm(const Gemm& g, hipStream_t st);

struct LibHandle;  // rocBLAS handle (below)

// ---------------------------------------------------------------- diabatic Ehrenfest
// A = (rho + rho^T)/2 with rho_ij = Re(ket_j,el bra_i,el) - Re(ket_j,hl bra_i,hl),
// then K = X o A - H' A.  bra and ket are row-major complex states (row r,
// column c at r * stride + c); conj_bra conjugates bra on load; bra_hl < 0
// leaves out the hole term.  X: ld x ld matrix (may alias K), or null to form
// X_ij from ip, kwh and vshift.  A and K: ld x ld, written in full.
struct EhrenfestArgs {
    int N, ld;
    const cplx* bra;
    long bra_stride;
    int bra_el, bra_hl, conj_bra;
    const cplx* ket;
    long ket_stride;
    int ket_el, ket_hl;
    const double* X;
    long ldx;
    const double *ip, *kwh, *vshift;
    double* A;
    double* K;
};
// lib: the product goes to rocBLAS dgemm (else the library's MFMA GEMM)
int launch_ehrenfest(const EhrenfestArgs& e, const double* Hp, LibHandle* lib, hipStream_t st);
// dst (ld x ld) = lower triangle of src, zero above the diagonal
void launch_copy_lower(const double* src, double* dst, int ld, hipStream_t st);

// ---------------------------------------------------------------- library calls (rocBLAS / rocSOLVER)
// One handle per context, on the context stream.
struct LibHandle;
LibHandle* lib_create(hipStream_t st);
void lib_destroy(LibHandle* es);
// C = alpha A B + beta C, column-major, no transposes
int lib_dgemm_nn(LibHandle* es, int m, int n, int k, double alpha, const double* A, int lda, const double* B, int ldb,
                 double beta, double* C, int ldc);
// The mixed-mode term products in S K-slices of KS: Y[z] = A[:, z KS : (z+1) KS] X[z KS : (z+1) KS, :]
// with A m x (S KS) column-major (FP32, or bf16 for the _bf16 form), X row-major with
// ldx values per row (the term's state, rows = orbitals), Y[z] column-major m x c
// (ldy) at stride c * ldy, FP32 products and accumulation within a slice.
int lib_gemm_slices_f32(LibHandle* es, int m, int c, int KS, int S, const float* A, int lda, const float* X, int ldx,
                        float* Y, int ldy);
int lib_gemm_slices_bf16(LibHandle* es, int m, int c, int KS, int S, const unsigned short* A, int lda,
                         const unsigned short* X, int ldx, float* Y, int ldy);
// Eigenpairs il..iu (1-based, ascending) of h C = S C E from H_orth and L^-1
// (Q2, lower triangle): E (device, 2 N doubles; eigenvalue j at E[j - 1]) and
// C = L^-T V (device, ld x nev); A and Z are ld x ld scratch.
int eigen_pairs(LibHandle* es, int N, int ld, const double* Hth, const double* Q2, double* A, double* Z, double* C,
                double* E, int il, int iu, int* d_ints, hipStream_t st);

// Cholesky S = L L^T in place (lower; upper zeroed) on an ld x ld padded
// matrix, plus L^-1 in Linv.  info (device int) gets the first failing
// column + 1, or stays 0.  Dinv: (ld/64) blocks of 64x64 workspace.
void launch_cholesky(double* A, int ld, double* Dinv, int* info, hipStream_t st);
// work: ld * ld / 2 doubles
void launch_trtri_lower(const double* L, double* Linv, const double* Dinv, int ld, double* work, hipStream_t st);
void launch_symmetrize_from_lower(double* A, int ld, hipStream_t st);
void launch_copy_matrix(const double* src, double* dst, long n, hipStream_t st);

// ---------------------------------------------------------------- propagation
enum Tri { TRI_FULL = 0, TRI_LOWER = 1, TRI_UPPER = 2 };
enum Epi { EPI_PLAIN = 0, EPI_CHEB_FIRST = 1, EPI_CHEB = 2 };

// Several time steps from one recurrence: accumulator q (q < m, layout of acc,
// at accm + q * accm_stride) is sum_k coefm[q * coefm_stride + k] T_k over
// k <= kq[q] (device table; kq[m - 1] is the series length).  The recurrence
// writes T_k (k >= 1) to ring + ((k - 1) % PANEL) * ring_stride; a panel pass
// adds terms k0..k1 (at most PANEL + 1 of them, k1 - k0 < PANEL after the
// first) to every accumulator that takes any of them, in term order.  k0 = 0:
// the pass starts the accumulators from zero and reads T_0 from phi.  n =
// elements of rows < N (row-major, R per row).
constexpr int MULTI_MAX = 10;  // accumulators of one multi-step series
constexpr int PANEL = 16;      // terms per panel pass, and the ring length
struct PanelArgs {
    cplx* accm;
    long accm_stride;
    const cplx* coefm;
    int coefm_stride;
    const cplx* ring;
    long ring_stride;
    const cplx* phi;
    const int* kq;
    int m;
    long n;
};
struct MatvecArgs {
    const double* M;  // column-major ld x ld; y_i = sum_j M[j + ld*i] x_j (= (M^T x)_i)
    int ld, N;
    int R;            // complex columns of x (row-major N x R)
    const cplx* x;    // T_{k-1} (or the input)
    const cplx* x2;   // T_{k-2}
    cplx* out;        // T_k or y
    cplx* acc;        // accumulator
    cplx* partial;    // [split][N][R]
    unsigned* counters;
    int split;
    int tri;
    int epi;
    double alpha, beta;     // T = alpha*y + beta*x - x2 (EPI_CHEB), alpha*y + beta*x (FIRST)
    const cplx* coef;       // device Chebyshev coefficients a_0..a_K (and a_{K+1} = 0)
    int k;                  // term index for EPI_CHEB
    int x_stride, x_off;    // column selection: uses x[i*x_stride + x_off + r], r < R
    int out_stride, out_off;
    const float* M32;       // FP32 copy of M (mixed and FP32-storage precision modes; may be null)
    const float* x32;       // mixed mode: FP32 copy of x (same layout), read by the FP32 GEMM; may be null
    float* out32;           // mixed mode: the epilogue also writes T_k here in FP32 (same layout); may be null
    unsigned short* outb16; // precision mode 3: the epilogue also writes T_k here in bf16 (same layout); may be null
    // A multi-step series' per-term accumulation (a null acc): with accm set,
    // accumulator q in [m_lo, m_hi) (layout of acc, at accm + q * accm_stride)
    // takes coefm[q * coefm_stride + k] T_k (EPI_CHEB), or is set to
    // coefm[q][0] x + coefm[q][1] T_1 (EPI_CHEB_FIRST).  Only the symmetric
    // block reduction (phase 2 of variants 16-18, 20, 21 and of the
    // FP32-stored operator's term) honours it and reports so through
    // take_multi_done; other terms ignore it, and their series add the terms
    // to the accumulators in panel passes instead.
    cplx* accm;
    long accm_stride;
    const cplx* coefm;
    int coefm_stride;
    int m_lo, m_hi;
    int m_pair;  // EPI_CHEB: 0 = adds nothing, 1 = c_k T_k, 2 = c_{k-1} x + c_k T_k (x = T_{k-1})
};
// true once after a term launch that accumulated a.accm itself
bool take_multi_done();
// A null acc leaves the accumulator out of the epilogue (out = T_k only): the
// recurrence of a multi-step series, whose accumulators take the terms in
// panels (launch_multi_panel).

void launch_multi_panel(const PanelArgs& p, int k0, int k1, hipStream_t st);
void launch_multi_panel_grouped(const PanelArgs* d_args, const int* task_desc, const int* task_blk, int ntask,
                                int k0, int k1, hipStream_t st);
int matvec_split_for(int N, int R);
// complex elements of a.partial the symmetric two-phase variant needs at (ld, R)
long sym_partial_cplx(int ld, int R);
long sym2_partial_cplx(int ld, int R);
void launch_strided_copy(cplx* dst, long ds, const cplx* src, long ss, int n, hipStream_t st);
bool sym_supported(int R, int tri);
void launch_matvec(const MatvecArgs& a, hipStream_t st);
void launch_matvec_variant(const MatvecArgs& a, int variant, hipStream_t st);
int matvec_variant_for(int N, int R, int tri);
void matvec_set_tuned(int N, int R, int tri, int variant);
bool matvec_is_tuned(int N, int R, int tri);
void matvec_force_variant(int v);
int matvec_variant_count();
void launch_cheb_gemm(const MatvecArgs& a, int prec, hipStream_t st);
// split-K matrix-core term GEMM (prec 0: FP64, 1: FP32 products on M32, 2: FP64
// products on M32 widened to FP64); P holds
// gemm3_partial_doubles(N, ld, R) doubles of partial products
void launch_gemm3(const MatvecArgs& a, int prec, double* P, hipStream_t st);
void launch_term_f32h(const MatvecArgs& a, hipStream_t st);
long gemm3_partial_doubles(int N, int ld, int R);
// Grouped column-block matvec over many systems (R in {1, 2, 4, 8, 16}); each
// task is (descriptor index, column block of grouped_cols_per_block(R) columns).
int grouped_cols_per_block(int R);
void launch_matvec_grouped(const MatvecArgs* d_args, const int* d_task_desc, const int* d_task_blk, int ntasks, int R,
                           int k, hipStream_t st);
// Symmetric half-storage term (4x4x4 products, R in {1, 2, 4, 8}) over many
// systems: phase-1 task i is segment t1_idx[i] of system t1_desc[i], phase-2
// task i is block * R + chunk = t2_idx[i] of system t2_desc[i]; sys_S[d] is
// system d's segment length in a batch, as sym_term_segments gives it.
int sym_term_segments(int ld, int* S);
// reduce_groups: the phase-2 group count of the batch's systems (sym_reduce_groups_for
// their ld; a batch whose systems differ in it uses the largest)
void launch_sym_grouped(const MatvecArgs* d_args, const int* t1_desc, const int* t1_idx, int n1, const int* t2_desc,
                        const int* t2_idx, int n2, const int* sys_S, int R, int k, int reduce_groups, hipStream_t st);
int sym_reduce_groups_for(int ld);

void launch_zero_u32(unsigned* p, int n, hipStream_t st);
void launch_zero_cplx(cplx* p, long n, hipStream_t st);
void launch_lanczos_start(int N, cplx* v, cplx* vprev, unsigned seed, hipStream_t st);
void launch_lanczos_step(int N, cplx* v, cplx* vprev, const cplx* w, double* alpha, double* beta, int j,
                         hipStream_t st);
// extreme-eigenvalue bounds of the Lanczos tridiagonal (al[0..m-1], be[1..m]),
// on the host: bounds = [emin, emax, theta_min, theta_max, res_min, res_max]
void lanczos_bounds_host(const double* al, const double* be, int m, double rel_margin, double* bounds);
int bessel_coef_host(double x, double phase, int kmax_alloc, double tol, cplx* coef);
void launch_bessel(double x, double phase, int kmax_alloc, double tol, double* J, cplx* coef, int* out_k,
                   hipStream_t st);
void launch_observables(int N, int W, int n_frag, const int* frag, const cplx* out, const cplx* phi, const cplx* hk,
                        double* pops, double* erg, double* norms, hipStream_t st);
void launch_column_norms(int N, int R, const cplx* x, double* out, hipStream_t st);
void launch_renorm_columns(cplx* x, int N, int R, const double* n0, const double* n1, hipStream_t st);
void launch_to_f32(const double* a, float* b, long n, hipStream_t st);
// H32 = fl32(H) and the bf16 residual H - H32, both column-major (n elements)
void launch_split_f32_bf16(const double* H, float* H32, unsigned short* Hlo, long n, hipStream_t st);
void launch_to_bf16(const float* a, unsigned short* b, long n, hipStream_t st);
// the mixed modes' library-GEMM slices added in FP64, then the term epilogue
void launch_lib_reduce(const MatvecArgs& a, const float* Phi, const float* Plo, int S, hipStream_t st);
void launch_identity_pad(double* S, double* h, int N, int ld, hipStream_t st);


// Small grouped helpers for multi-system batches.
struct NormTask {
    const cplx* x;
    int N, R, c;
    double* out;
};
struct CopyTask {
    const cplx* src;
    cplx* dst;
    long n;
};
struct ObsTask {
    int N, W, n_frag, w;
    const int* frag;
    const cplx* out;
    const cplx* phi;
    const cplx* hk;
    double *pops, *erg, *norms;
};
void launch_grouped_norms(const NormTask* d_tasks, int n, hipStream_t st);
void launch_grouped_copy(const CopyTask* d_tasks, int n, hipStream_t st);
void launch_grouped_observables(const ObsTask* d_tasks, int n, hipStream_t st);

}  // namespace dh
