/*
 * dynemol_hip: MI355X (gfx950) electron/hole wavepacket propagator.
 *
 * C ABI: plain pointers and sizes, int status returns (0 = success,
 * negative = error; dh_last_error() describes the last failure of the
 * calling thread's most recent call).  Matrices are column-major (Fortran
 * order); complex vectors are interleaved (re, im) double pairs, i.e.
 * Fortran complex*16.  Host pointers are plain pageable or pinned memory.
 *
 * Threading: a context owns one HIP stream; calls on one context must not
 * overlap.  Systems created on a context run on that context's stream.
 * Several contexts may be driven from several host threads at once; the
 * term series' stream capture (dh_set_graph, dh_batch_set_graph; on by
 * default for N <= 2048) must then be off, since a capture must not overlap
 * another thread's device allocations and frees.
 */
#ifndef DYNEMOL_HIP_H
#define DYNEMOL_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    DH_OK = 0,
    DH_ERR_ARG = -1,      /* invalid argument or unsupported basis */
    DH_ERR_HIP = -2,      /* HIP runtime error */
    DH_ERR_NOT_SPD = -3,  /* Cholesky of S failed: S not positive definite */
    DH_ERR_DIVERGED = -4, /* Chebyshev series lost the norm after widening the bounds */
    DH_ERR_NOMEM = -5,    /* device allocation failed */
    DH_ERR_STATE = -6     /* call out of order (e.g. propagate before operators) */
};

enum { DH_MAT_S = 0, DH_MAT_H = 1, DH_MAT_L = 2, DH_MAT_LINV = 3, DH_MAT_HORTH = 4, DH_MAT_HPRIME = 5 };

typedef struct dh_context dh_context;
typedef struct dh_system dh_system;

/* Message of the last failed call on this thread ("" if none). */
const char *dh_last_error(void);
int dh_version(void);

/* Replaces gpu_init_ (GPU_Interface.cpp:226-263): select the device, create
 * the stream.  The reference picks the device round-robin by MPI rank; here
 * the caller passes it. */
int dh_init(int device, dh_context **ctx);
/* Replaces gpu_finalize_ (GPU_Interface.cpp:265-285). */
int dh_finalize(dh_context *ctx);

/*
 * A system: device-resident operators and wavepackets of one molecule.
 * The basis arguments are the reference's STO_basis record in struct-of-
 * arrays form (types_EHT.f:114-166; built by Basis_Builder, structure.f:
 * 228-324): per orbital n, l, m, Nzeta, zeta(2) [1/Angstrom], coef(2), IP
 * [eV], k_WH, V_shift, fragment index; per atom the first orbital
 * (BasisPointer) and AngMax.  n_atoms = 0 creates an operator-only system
 * (S and h loaded from the host).  n_wp wavepackets (electron, hole, ...)
 * each carry a bra and a ket (Psi_t_bra / Psi_t_ket, ElHl_Chebyshev.f:40).
 */
int dh_system_create(dh_context *ctx, int n_atoms, const int *first_orb, const int *atom_lmax, int n_orb,
                     const int *orb_n, const int *orb_l, const int *orb_m, const int *orb_nzeta,
                     const double *orb_zeta, const double *orb_coef, const double *orb_ip, const double *orb_kwh,
                     const double *orb_vshift, const int *orb_frag, int n_frag, int n_wp, dh_system **sys);
int dh_system_destroy(dh_system *sys);
/* Device bytes a system of this size holds (serving admission), in FP64 and
 * with the FP32 operator copies of a precision mode (dh_set_precision);
 * dh_system_bytes_prec(n, w, 0) also counts the accumulators of the
 * multi-step series of dh_propagate_steps. */
long long dh_system_bytes(int n_orb, int n_wp);
long long dh_system_bytes_prec(int n_orb, int n_wp, int precision);

/* Atomic coordinates, n_atoms x 3 row-major (x, y, z per atom), Angstrom
 * (system%coord, types_EHT.f).  */
int dh_set_coords(dh_system *sys, const double *coords);
/* Overlap S and Hueckel h from the coordinates, on the device
 * (Overlap_Matrix overlap_D.f:31-106 + Build_Huckel ElHl_Chebyshev.f:296-323),
 * then dh_orthogonalize.  cutoff: QM_cutoff (12 A, constants_m.f:34).
 * s_drop > 0 zeroes off-diagonal |S_ij| < s_drop as the reference's
 * unmoved-atom recycle cache does (overlap_D.f:193-225, 586-655). */
int dh_build_operators(dh_system *sys, double cutoff, double s_drop);
/* Host S and h (N x N, leading dimension ld), then dh_orthogonalize: the
 * operands propagationelhl_gpucaller_ receives (Taylor_gpu.cpp:634-682). */
int dh_load_operators(dh_system *sys, const double *S, const double *h, int ld);
/* S = L L^T, L^-1, H_orth = L^-1 h L^-T, Lanczos spectral bounds; with
 * want_hprime also H' = S^-1 h (replaces ElHl_Chebyshev.f:205-210 and
 * GPU_Interface.cpp:880-949 + Taylor_gpu.cpp:687-696). */
int dh_orthogonalize(dh_system *sys, int want_hprime);
/* Copy a device matrix (DH_MAT_*) to host, N x N with leading dimension ld.
 * DH_MAT_L and DH_MAT_LINV come back mirrored: the lower triangle holds the
 * factor and the upper triangle its transpose (the storage the triangular
 * products read), so take the lower triangle for L and L^-1. */
int dh_get_matrix(dh_system *sys, int which, double *out, int ld);
/* [E_min, E_max, theta_min, theta_max, residual_min, residual_max] (eV). */
int dh_get_bounds(dh_system *sys, double *out6);
/* Override the Lanczos step count and relative safety margin (defaults 48, 0.01). */
int dh_set_lanczos(dh_system *sys, int steps, double rel_margin);

/* Arithmetic of the Chebyshev term products: 0 = FP64 (default); 1 = mixed:
 * FP32 matrix-core products with FP32 accumulation on an FP32 copy of H_orth,
 * while the recurrence, Bessel coefficients, accumulators, norms and
 * observables stay FP64; 2 = FP32-stored operator: the terms read an FP32
 * copy of H_orth (half the bytes of the HBM-bound few-wavepacket term) and
 * widen it to FP64, so products and sums are FP64 and the only error is the
 * rounding of H_orth itself; 3 = mixed with the FP64 residual correction: mode
 * 1's FP32 products plus the products of the residual H_orth - fl32(H_orth)
 * (stored in bf16, bf16 matrix-core products with FP32 accumulators of their
 * own, added to the main products in FP64), so the terms apply H_orth to 2^-32 relative instead of
 * 2^-24.  Modes 1 and 3 restore each step's FP64 norms.  No counterpart in the
 * reference (all FP64). */
int dh_set_precision(dh_system *sys, int mode);

/* Wavepacket wp's AO-basis Psi_t_bra and Psi_t_ket, N complex each. */
int dh_set_wavepacket(dh_system *sys, int wp, const double *psi_bra, const double *psi_ket);
int dh_get_wavepacket(dh_system *sys, int wp, double *psi_bra, double *psi_ket);
/* Initial state from AO coefficients (the FMO/eigenvector C the caller
 * chose): Psi_ket = C_ket, Psi_bra = S C_bra, as preprocess_ElHl_Chebyshev
 * forms them (ElHl_Chebyshev.f:110-129); also computes the t = 0 observables.
 * Needs built operators. */
int dh_set_initial_state(dh_system *sys, int wp, const double *ao_bra, const double *ao_ket);
/* Eigenpairs il..iu (1-based, ascending) of the generalised problem
 * h C = S C E of the system's current operators, computed on the device
 * (eigenpairs of H_orth, then C = L^-T V, so C^T S C = I): E (iu - il + 1,
 * eV) and C (N x (iu - il + 1), column-major, ldc >= N); either may be NULL.
 * The sign of each column is arbitrary, as with any eigensolver.  Replaces
 * the dsygvd solve of FMO (FMO.f:294-350; xPU_dsygvd, GPU_Interface.cpp:705-853)
 * for whole-system orbitals. */
int dh_eigenpairs(dh_system *sys, int il, int iu, double *E, double *C, int ldc);
/* Seed wavepacket wp with eigenvector k (1-based) of (h, S): AO_bra = AO_ket
 * = C_k, as the electron (LUMO) and hole (HOMO) start when all atoms belong to
 * both fragments (tuning.f:76-83, ElHl_Chebyshev.f:110-129).  energy (may be
 * NULL) receives E_k.  Then as dh_set_initial_state. */
int dh_set_eigenstate(dh_system *sys, int wp, int k, double *energy);
/*
 * Propagate every wavepacket from t_init to t_max (ps): Psi_ket <- exp(-i
 * S^-1 h dt / hbar) Psi_ket and Psi_bra <- exp(-i h S^-1 dt / hbar) Psi_bra,
 * the operators Propagation applies (Taylor.f:35-127).  n_terms (may be
 * NULL) receives the Chebyshev term count used.
 */
int dh_propagate(dh_system *sys, double t_init, double t_max, int *n_terms);
/*
 * n_steps nuclear steps in one call, the loop of the driver (Chebyshev_driver.f:
 * 102-169 calling ElHl_Chebyshev, ElHl_Chebyshev.f:180-283) with the state, the
 * operators and the per-step outputs kept on the device: each step propagates
 * every wavepacket by dt (ps) as dh_propagate does; with rebuild_every = r > 0
 * the operators are rebuilt (S, h, Cholesky, H_orth, bounds) before steps r,
 * 2r, ... from geometry g = step / r - 1 of coords (n_atoms x 3 per geometry,
 * (n_steps - 1) / r geometries, all copied to the device at the start).  Each
 * rebuild runs on a second stream of the context, into a second set of
 * operator buffers, while the interval before it propagates (it is launched
 * before that interval's steps, without host waits), and is swapped in at its
 * step (the system and the context then hold about 7 N^2 doubles more).  The
 * host waits once per rebuild interval, where the norms of the interval's
 * steps and the rebuild's Cholesky status are checked (a failed rebuild
 * returns DH_ERR_NOT_SPD); an interval with a drifting step is redone step by step
 * with wider bounds.  A rebuild takes the context workspace over, as
 * dh_build_operators does.  pops (n_steps x n_wp x (n_frag + 1)) and erg (n_steps x
 * n_wp x 2) receive each step's populations and energies (may be NULL); the
 * final state is read as after dh_propagate.  With dh_set_steps_per_series(sys,
 * 1): the same kernels and results as n_steps calls of dh_propagate(sys, 0, dt)
 * with dh_set_coords + dh_build_operators at the rebuild steps; by default FP64
 * steps share recurrences (dh_set_steps_per_series) and agree with those to
 * rounding.
 */
int dh_propagate_steps(dh_system *sys, int n_steps, double dt, int rebuild_every, const double *coords,
                       double cutoff, double s_drop, double *pops, double *erg, int *n_terms);
/* Observables after dh_propagate for wavepacket wp:
 * ao_bra (N complex) = S^-1 Psi_t_bra, not conjugated (Taylor_gpu.cpp:716-718);
 * pops (n_frag + 1) = Populations: Re sum_frag conj(Psi_ket) Psi_bra, total
 * last (data_output.f:87-147); erg (2) = QuasiParticleEnergies
 * sum_ij conj(ao_bra)_i h_ij Psi_ket_j (ElHl_Chebyshev.f:329-371). */
int dh_get_observables(dh_system *sys, int wp, double *ao_bra, double *pops, double *erg);
/* Diabatic-Ehrenfest kernel matrix of the current state, on the device:
 * K = X o A - H' A (N x N, column-major, ldk >= N, written to host K), with
 * A = (rho + rho^T) / 2 and rho_ij = Re(AO_ket_j,el AO_bra_i,el)
 * - Re(AO_ket_j,hl AO_bra_i,hl), where AO_bra = conj(S^-1 Psi_bra) and
 * AO_ket = Psi_ket of wavepackets wp_el and wp_hl (wp_hl = -1: no hole term).
 * X (N x N, ldx >= N) holds the extended-Hueckel factors h_ij = X_ij S_ij, or is
 * NULL to form them from the system's basis (systems built from coordinates).
 * H' is formed on the device on the first call after each operator build
 * (from L^-1 h, which the build leaves in the context) and stays there.  K may be
 * NULL: the matrix is then only computed (it stays in the context until the next
 * call).  A later build or load of another system on the same context takes
 * the workspace: the call then fails with DH_ERR_STATE until this system is
 * built again.  Replaces ehrenfestkernel2_gpu_ (Taylor_gpu.cpp:800-875) with the
 * calculate_rho / Huckel_stuff steps before it (diabatic-Ehren.f:100-116,
 * 250-276, 366-392); the host copies of H', A and X it needed are gone. */
int dh_ehrenfest_kernel(dh_system *sys, int wp_el, int wp_hl, const double *X, int ldx, double *K, int ldk);
/* Counters (9 doubles): [terms of the last recurrence (one step's, or a
 * multi-step series' in dh_propagate_steps), steps, matvec launches, divergence
 * retries, last norm drift, term-series GPU ms (HIP events, when timing is
 * on), term launches timed, operator rebuilds, host ms that dh_propagate_steps
 * spent launching its rebuilds beside the propagation and waiting for them]. */
int dh_get_stats(dh_system *sys, double *out9);
/* Record HIP events around each step's Chebyshev term series (on the
 * context stream that runs them); no counterpart in the reference, whose
 * only timing hook is the GPU_TIMING wall-clock print (GPU_Interface.cpp:33-49). */
int dh_set_timing(dh_system *sys, int on);
/* Steps per Chebyshev recurrence in dh_propagate_steps (every precision mode
 * and term kernel): within a rebuild interval the operator is fixed and the
 * terms T_k(Hh) phi do not depend on time, so one recurrence to the last
 * step's K collects every step's state, each with its own coefficients (the
 * symmetric half-storage terms add each term to the step accumulators in
 * their epilogue; the other term kernels leave the terms in a 16-slot ring
 * that a pass adds to the accumulators every 16 terms).  m = 0
 * (default): up to 10 steps per recurrence, within the interval and the
 * 4096-term coefficient table; m = 1: one recurrence per step, the same
 * kernels and bits as dh_propagate.  The steps' states agree to rounding
 * (the propagator is exact either way) with fewer terms: at N = 4096 (TiO2
 * slab) 10 steps take about 2110 terms instead of 10 x 261.  No counterpart
 * in the reference (one Taylor expansion per step). */
int dh_set_steps_per_series(dh_system *sys, int m);
/* Run each step's Chebyshev term series as one captured hipGraph (on = 1;
 * the default for N <= 2048) or as individual kernel launches (on = 0).  The graph is
 * captured on the first step with given spectral bounds, term count and buffer
 * assignment and replayed afterwards; results are identical either way.  No
 * counterpart in the reference (one cuBLAS call per term, Taylor_gpu.cpp:295-330). */
int dh_set_graph(dh_system *sys, int on);
/* Matvec kernel variant chosen by the autotuner for this system's shapes:
 * [Chebyshev term (2*n_wp RHS), <H> (n_wp), L phi_bra, L^-T phi, L^T ket,
 * L^-1 bra, Lanczos (1 RHS)].  No counterpart in the reference (its GEMV tile
 * sizes are compile-time macros, dzgemv_kernels.cu:581-608). */
int dh_matvec_tuning(dh_system *sys, int *out7);

/*
 * Multi-system batch: n systems of possibly different N with the same number
 * of wavepackets (1, 2, 4, 8 or 16), advanced together by one grouped launch per
 * Chebyshev term (FP64 members; members in other precision modes run their own
 * term kernels in the same step).  The systems stay owned by the caller; results are read per
 * system with dh_get_wavepacket / dh_get_observables.  No counterpart in the
 * reference (one system per process, ElHl_Chebyshev.f:148-291).
 */
typedef struct dh_batch dh_batch;
int dh_batch_create(dh_context *ctx, int n, dh_system *const *systems, dh_batch **batch);
int dh_batch_destroy(dh_batch *batch);
int dh_batch_propagate(dh_batch *batch, double t_init, double t_max, int *max_terms);
/* n_steps steps of dt (ps) for every member, up to 10 steps per Chebyshev
 * recurrence (each member's steps collected from the same terms with their
 * own coefficients, as dh_propagate_steps does for one system): one norm
 * check, commit and output pass per recurrence, and the outputs read
 * afterwards are those of the final state.  The states agree with n_steps
 * calls of dh_batch_propagate to rounding.  The grouped terms write each
 * member's ring of 16 terms, and a grouped pass adds them to the members'
 * accumulators every 16 terms.  Batches with members in other precision modes
 * step one step at a time.  max_terms: the most terms of one recurrence.  No counterpart
 * in the reference. */
int dh_batch_propagate_steps(dh_batch *batch, int n_steps, double dt, int *max_terms);
/* Term kernel family the batch autotuner chose for this batch's wavepacket
 * count (once per process, by timing both on the first batch of that count):
 * 1 = symmetric half-storage tiles, 0 = dense column blocks.  Members in a
 * precision mode other than FP64 run their own term kernels inside the batch
 * step.  No counterpart in the reference. */
int dh_batch_tuning(dh_batch *batch, int *family);
/* Run each step's Chebyshev term launches as one captured hipGraph (on = 1,
 * the default; captured again when a system's term count changes) or as
 * individual launches (on = 0).  Results are identical either way.  No
 * counterpart in the reference. */
int dh_batch_set_graph(dh_batch *batch, int on);

/*
 * Drop-in for propagationelhl_gpucaller_ (Taylor_gpu.cpp:219-232, 634-736):
 * host S, h (N x N, lda N) -> H_out = S^-1 h (may be NULL), PSI_bra/PSI_ket
 * propagated in place from *t_init to *t_max, AO_bra = S^-1 PSI_bra (not
 * conjugated), AO_ket unused, *tau and *save_tau set to the step in eV^-1.
 * All arguments by reference like the reference's implicit Fortran interface.
 */
int dh_propagation_elhl(dh_context *ctx, const int *N, const double *S, const double *h, double *H_out,
                        double *AO_bra, double *AO_ket, double *PSI_bra, double *PSI_ket, const double *t_init,
                        const double *t_max, double *tau, double *save_tau);

/*
 * Propagation of caller-owned device operands: phi <- exp(-i H dt / hbar) phi
 * for R (1, 2, 4 or 8) complex columns, by the Chebyshev series on
 * [bounds[0], bounds[1]] (eV; e.g. dh_get_bounds) with the same term kernels
 * and norm check as dh_propagate (the role of Propagation, Taylor.f:35-127, on
 * an already orthogonalised operator; Taylor_gpu.cpp:295-330 applies H' the
 * same way once per term).  H: device, ldh x ldh column-major, symmetric, ldh =
 * N rounded up to a multiple of 64, rows and columns N..ldh-1 zero off the
 * diagonal (as H_orth is stored).  phi: device, ldh rows x R complex, row-major
 * (column c of row i at phi[2 (i R + c)]); rows N..ldh-1 are set to zero.
 * dt in ps.  work: a 256-byte aligned device buffer of at least
 * dh_chebyshev_work_bytes(N, R) bytes.  Runs on the context stream and returns
 * when phi holds the result; n_terms (may be NULL) receives the term count.
 */
long long dh_chebyshev_work_bytes(int N, int R);
int dh_chebyshev_apply(dh_context *ctx, int N, const double *H, int ldh, const double *bounds, double dt, int R,
                       double *phi, void *work, long long work_bytes, int *n_terms);
/*
 * dh_chebyshev_apply without the work argument and without any device
 * allocation: for callers that cannot allocate (code under stream capture set
 * up beforehand, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  The work buffer is a static 12 MiB device
 * array, one per GPU, held until the process ends: N <= 4096 with R <= 4, or
 * N <= 2048 with R <= 8 (dh_chebyshev_work_bytes(N, R) <= 12 MiB); larger
 * shapes are refused with DH_ERR_ARG.  Every call on a GPU shares that GPU's
 * array, so calls that overlap on one GPU (from several contexts or threads)
 * race.  Same kernels and results as dh_chebyshev_apply.
 */
int dh_chebyshev_apply_noalloc(dh_context *ctx, int N, const double *H, int ldh, const double *bounds, double dt,
                               int R, double *phi, int *n_terms);

#ifdef __cplusplus
}
#endif
#endif
