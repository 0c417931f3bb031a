// C ABI of the dynemol_hip library: contexts, systems and the drop-in
// propagationelhl_gpucaller_ replacement.  Host orchestration only; every
// arithmetic step runs in the kernels of overlap.hip, dense.hip and cheb.hip.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/dynemol_hip.h"
#include "common.hpp"
#include "kernels.hpp"

namespace dh {
static thread_local char g_err[512] = "";
void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}
const char* last_error() { return g_err; }
}  // namespace dh

using namespace dh;

static const double H_BAR = 6.58264e-4;  // eV*ps, constants_m.f:23
static const int KMAX_ALLOC = 4096;      // Chebyshev terms supported (r*dt up to ~3500)
static const int LANCZOS_MAX = 120;
// Chebyshev series truncation: the last term kept has |J_k(r dt)| above this.
// The neglected tail is below 2.2e-16 of the state's norm, under the FP64
// rounding the ~10^2 terms of a step accumulate.
static const double BESSEL_TOL = 1e-16;
// K slices of the mixed modes' library GEMMs (ld is a multiple of 64, so the
// slices are equal and whole)
static const int LIB_SLICES = 4;

// Scratch of dh_chebyshev_apply_noalloc: one copy per device, held until the
// process ends.  12 MiB covers the work layout of N <= 4096 with R <= 4 and of
// N <= 2048 with R <= 8.  Host code never uses the array's name as an address:
// dh_init looks up each device's copy with hipGetSymbolAddress and keeps it in
// g_scratch_dev.
static constexpr long long NOALLOC_SCRATCH_BYTES = 12LL << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
static constexpr int MAX_DEVICES = 64;
static void* g_scratch_dev[MAX_DEVICES] = {};
static std::mutex g_scratch_mu;

// Device address of `scratch` on the current device (once per device).
static int scratch_setup(int device) {
    if (device < 0 || device >= MAX_DEVICES) {
        dh::set_error("dh_init: device index %d beyond %d", device, MAX_DEVICES);
        return DH_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_dev[device]) return DH_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        dh::set_error("dh_init: hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
        return DH_ERR_HIP;
    }
    g_scratch_dev[device] = ptr;
    return DH_OK;
}

struct dh_context {
    int device = 0;
    hipStream_t stream = nullptr;
    // shared workspace for operator builds (one build at a time per context)
    int ws_ld = 0;
    double *S = nullptr, *h = nullptr, *W = nullptr, *Hp = nullptr, *Dinv = nullptr;
    double* Kd = nullptr;  // Ehrenfest kernel matrix / eigensolver scratch (allocated on first use)
    LibHandle* lib = nullptr;  // rocBLAS handle on the context stream (created on first use)
    int* d_info = nullptr;
    const dh_system* ws_owner = nullptr;  // system whose S and h the workspace holds
    dh_system* dropin = nullptr;
    int dropin_N = 0;
    // operator builds overlapped with propagation (dh_propagate_steps): a
    // second stream and a second workspace of ws_ld, swapped with the first
    // when the build is taken over
    hipStream_t side = nullptr;
    double *S2 = nullptr, *h2 = nullptr, *W2 = nullptr, *Hp2 = nullptr, *Dinv2 = nullptr;
    int* info2 = nullptr;
    int ws2_ld = 0;
};

struct dh_system {
    dh_context* ctx = nullptr;
    int n_atoms = 0, N = 0, ld = 0, W = 0, n_frag = 1;
    // basis
    BasisDev bd{};
    void* basis_mem = nullptr;
    double* coords = nullptr;
    int* frag = nullptr;
    // operators
    double *Q1 = nullptr, *Q2 = nullptr, *Hth = nullptr;
    // propagation buffers
    cplx *phi = nullptr, *x0 = nullptr, *x1 = nullptr, *acc = nullptr, *out = nullptr, *hk = nullptr;
    cplx* partial = nullptr;
    cplx* stage = nullptr;  // 2 x ld contiguous columns for host transfers
    unsigned* counters = nullptr;
    int split_max = 1;
    // Lanczos / Bessel
    cplx *lv = nullptr, *lvp = nullptr, *lw = nullptr;
    double *lal = nullptr, *lbe = nullptr, *J = nullptr;
    cplx* coef = nullptr;
    int* kout = nullptr;
    cplx* coef_h = nullptr;         // pinned staging of the coefficients
    hipEvent_t coef_ev = nullptr;   // the last copy out of coef_h
    bool coef_pending = false;
    // multi-step series of dh_propagate_steps: MULTI_MAX accumulators and their
    // coefficient rows (KMAX_ALLOC + 2 each; device, and pinned staging), and
    // the PANEL-slot ring of the recurrence's terms
    int steps_per_series = 0;       // 0: as many as the interval allows; 1: one series per step
    cplx *accm = nullptr, *coefm = nullptr, *coefm_h = nullptr, *ring = nullptr;
    int *kq_d = nullptr, *kq_h = nullptr;  // the accumulators' last terms (device; pinned staging)
    hipEvent_t coefm_ev = nullptr;
    bool coefm_pending = false;
    // observables
    double *pops = nullptr, *erg = nullptr, *norms = nullptr, *norms0 = nullptr, *norms1 = nullptr;
    // host state
    bool have_coords = false, ops_ready = false, phi_dirty = true, state_set = false, hprime_ready = false;
    bool w_ready = false;  // the context's W holds L^-1 h of this system's current operators
    double bounds_h[6] = {0, 0, 0, 0, 0, 0};
    double emin = 0, emax = 0;
    int lanczos_steps = 48;
    double rel_margin = 0.01;  // of the Ritz width on each side, beyond the Ritz residuals
    double widen = 1.0;
    double coef_dt = -1.0, coef_c = 0.0, coef_r = 0.0;
    int K = 0;
    // [0] terms of last step, [1] steps, [2] matvec launches, [3] divergence retries,
    // [4] last norm drift, [5] term-series GPU ms (events), [6] term launches timed,
    // [7] operator rebuilds, [8] rebuild wall ms
    double stats[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    bool timing = false;
    // captured term series (hipGraph) for the two phi/acc buffer assignments;
    // the key is everything the captured kernel arguments depend on
    struct SeriesGraph {
        hipGraphExec_t exec = nullptr;
        int K = -1, prec = -1, variant = -1;
        const void* phi = nullptr;
        const void* H = nullptr;  // the operator buffer (swapped by overlapped rebuilds)
        double c = 0.0, r = 0.0;
    } graphs[2];
    int next_graph = 0;
    bool use_graph = true;
    bool tuned = false;
    // 0 = FP64; 1 = mixed (FP32 MFMA term GEMM); 2 = FP32-stored H_orth with
    // FP64 products and sums
    int precision = 0;
    float* H32 = nullptr;     // FP32 copy of H_orth for modes 1 and 2
    // mode 1: FP32 copies of the term inputs (the step's initial state, then
    // each T_k as its epilogue writes it), alternating between two buffers
    float* x32[2] = {nullptr, nullptr};
    bool h32_ready = false;
    unsigned short* Hlo = nullptr;  // mode 3: bf16 residual H_orth - H32, column-major
    // mode 3: bf16 copies of the term inputs (the residual's products), alternating like x32
    unsigned short* xb16[2] = {nullptr, nullptr};
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<hipEvent_t> evpool;  // per-step series events of dh_propagate_steps (timing on)
    // the next geometry's operators, built on the context's side stream while
    // the current interval's steps run (dh_propagate_steps), then swapped in
    struct Shadow {
        double *Q1 = nullptr, *Q2 = nullptr, *Hth = nullptr, *coords = nullptr;
        float* H32 = nullptr;
        unsigned short* Hlo = nullptr;
        cplx *lv = nullptr, *lvp = nullptr, *lw = nullptr, *partial = nullptr;
        double *lal = nullptr, *lbe = nullptr;
        unsigned* counters = nullptr;
        double bounds[6] = {0, 0, 0, 0, 0, 0};
        // a launched build's Cholesky status and Lanczos coefficients come back
        // into pinned host memory; `done` follows them on the side stream
        int* h_info = nullptr;
        double *h_al = nullptr, *h_be = nullptr;
        hipEvent_t done = nullptr;
        int m = 0;
        bool pending = false;
        bool allocated = false;
    } sh;
};

// modes whose term products are FP32 (1, and 3 with its residual correction):
// their FP32 rounding drifts the norm ~1e-7 per step, which a per-step FP64
// norm restoration removes
static bool fp32_products(const dh_system* s) { return s->precision == 1 || s->precision == 3; }

#define CHECK_ARG(cond, msg)                  \
    do {                                      \
        if (!(cond)) {                        \
            dh::set_error("%s", msg);         \
            return DH_ERR_ARG;                \
        }                                     \
    } while (0)

template <class T>
static int dalloc(T** p, size_t count) {
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)p, count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        dh::set_error("hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
        return DH_ERR_NOMEM;
    }
    return DH_OK;
}
#define DALLOC(p, n)                    \
    do {                                \
        int rc_ = dalloc(&(p), (n));    \
        if (rc_) return rc_;            \
    } while (0)

static int round_ld(int N) { return ((N + TILE - 1) / TILE) * TILE; }

extern "C" const char* dh_last_error(void) { return dh::last_error(); }
extern "C" int dh_version(void) { return 100; }

extern "C" int dh_init(int device, dh_context** out) {
    CHECK_ARG(out != nullptr, "dh_init: null output");
    int n = 0;
    DH_HIP(hipGetDeviceCount(&n));
    CHECK_ARG(device >= 0 && device < n, "dh_init: device index out of range");
    DH_HIP(hipSetDevice(device));
    const int src = scratch_setup(device);
    if (src) return src;
    dh_context* c = new dh_context();
    c->device = device;
    DH_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    if (dalloc(&c->d_info, 1)) {
        delete c;
        return DH_ERR_NOMEM;
    }
    *out = c;
    return DH_OK;
}

static void free_ws2(dh_context* c) {
    if (c->side) hipStreamSynchronize(c->side);
    for (double* p : {c->S2, c->h2, c->W2, c->Hp2, c->Dinv2})
        if (p) hipFree(p);
    if (c->info2) hipFree(c->info2);
    c->S2 = c->h2 = c->W2 = c->Hp2 = c->Dinv2 = nullptr;
    c->info2 = nullptr;
    c->ws2_ld = 0;
}

static void free_ws(dh_context* c) {
    hipFree(c->S); hipFree(c->h); hipFree(c->W); hipFree(c->Hp); hipFree(c->Dinv);
    if (c->Kd) hipFree(c->Kd);
    c->S = c->h = c->W = c->Hp = c->Dinv = c->Kd = nullptr;
    c->ws_ld = 0;
    c->ws_owner = nullptr;
    free_ws2(c);
}

static int ensure_ws(dh_context* c, int ld) {
    if (c->ws_ld >= ld) return DH_OK;
    DH_HIP(hipStreamSynchronize(c->stream));
    free_ws(c);
    const size_t n = (size_t)ld * ld;
    DALLOC(c->S, n);
    DALLOC(c->h, n);
    DALLOC(c->W, n);
    DALLOC(c->Hp, n);
    DALLOC(c->Dinv, (size_t)ld * TILE);
    c->ws_ld = ld;
    return DH_OK;
}

extern "C" int dh_system_destroy(dh_system* s) {
    if (!s) return DH_OK;
    if (s->ctx->ws_owner == s) s->ctx->ws_owner = nullptr;
    hipSetDevice(s->ctx->device);
    hipStreamSynchronize(s->ctx->stream);
    if (s->H32) hipFree(s->H32);
    if (s->Hlo) hipFree(s->Hlo);
    for (float* p : s->x32)
        if (p) hipFree(p);
    for (unsigned short* p : s->xb16)
        if (p) hipFree(p);
    for (auto& g : s->graphs)
        if (g.exec) hipGraphExecDestroy(g.exec);
    if (s->ev0) hipEventDestroy(s->ev0);
    if (s->ev1) hipEventDestroy(s->ev1);
    if (s->coef_ev) hipEventDestroy(s->coef_ev);
    if (s->coef_h) hipHostFree(s->coef_h);
    if (s->coefm_ev) hipEventDestroy(s->coefm_ev);
    if (s->coefm_h) hipHostFree(s->coefm_h);
    if (s->kq_h) hipHostFree(s->kq_h);
    if (s->accm) hipFree(s->accm);
    if (s->ring) hipFree(s->ring);
    if (s->coefm) hipFree(s->coefm);
    if (s->kq_d) hipFree(s->kq_d);
    for (hipEvent_t e : s->evpool) hipEventDestroy(e);
    if (s->ctx->side) hipStreamSynchronize(s->ctx->side);  // a launched build may still run
    void* shp[] = {s->sh.Q1, s->sh.Q2, s->sh.Hth, s->sh.coords, s->sh.H32, s->sh.Hlo, s->sh.lv, s->sh.lvp, s->sh.lw,
                   s->sh.partial, s->sh.lal, s->sh.lbe, s->sh.counters};
    for (void* p : shp)
        if (p) hipFree(p);
    for (void* p : {(void*)s->sh.h_info, (void*)s->sh.h_al, (void*)s->sh.h_be})
        if (p) hipHostFree(p);
    if (s->sh.done) hipEventDestroy(s->sh.done);
    void* ptrs[] = {s->basis_mem, s->coords, s->frag, s->Q1, s->Q2, s->Hth, s->phi, s->x0, s->x1, s->acc, s->out, s->stage,
                    s->hk, s->partial, s->counters, s->lv, s->lvp, s->lw, s->lal, s->lbe, s->J, s->coef,
                    s->kout, s->pops, s->erg, s->norms, s->norms0, s->norms1};
    for (void* p : ptrs)
        if (p) hipFree(p);
    delete s;
    return DH_OK;
}

extern "C" int dh_finalize(dh_context* c) {
    if (!c) return DH_OK;
    hipSetDevice(c->device);
    if (c->dropin) dh_system_destroy(c->dropin);
    hipStreamSynchronize(c->stream);
    free_ws(c);
    lib_destroy(c->lib);
    hipFree(c->d_info);
    if (c->side) hipStreamDestroy(c->side);
    hipStreamDestroy(c->stream);
    delete c;
    return DH_OK;
}

// Partial-sum buffer of a system: split-K variants (up to 16 RHS per pass) and
// the symmetric two-phase variant for the RHS counts a system uses.
static long partial_cplx_for(int N, int n_wp) {
    const int ld = round_ld(N);
    long n = (long)matvec_split_for(N, 1) * ld * 16;
    const int rs[3] = {2 * n_wp, n_wp, 1};
    for (int R : rs) {
        if (sym_supported(R, TRI_FULL)) n = std::max(n, std::max(sym_partial_cplx(ld, R), sym2_partial_cplx(ld, R)));
        n = std::max(n, (gemm3_partial_doubles(N, ld, R) + 1) / 2);
    }
    // the mixed modes' library GEMM slices: FP32 and residual, LIB_SLICES x 2R x ld floats each
    n = std::max(n, (long)LIB_SLICES * 2 * (2 * n_wp) * ld * 2 / 4);
    return n;
}

extern "C" long long dh_system_bytes(int n_orb, int n_wp) {
    const long long ld = round_ld(n_orb);
    const long long R = 2LL * n_wp;
    long long b = 3LL * ld * ld * 8;                       // Q1, Q2, H_orth
    b += (4 * R + 3LL * n_wp + n_wp) * ld * 16;             // phi, x0, x1, acc, out, hk
    b += (long long)partial_cplx_for(n_orb, n_wp) * 16;    // partials
    b += 3LL * ld * 16 + (long long)KMAX_ALLOC * 24;      // Lanczos, Bessel
    return b;
}

extern "C" long long dh_system_bytes_prec(int n_orb, int n_wp, int precision) {
    const long long ld = round_ld(n_orb);
    long long b = dh_system_bytes(n_orb, n_wp);
    if (precision != 0) b += ld * ld * 4;                       // H32
    if (precision == 3) b += ld * ld * 2;                       // bf16 residual
    if (precision == 1 || precision == 3) b += 2LL * ld * 4 * n_wp * 4;  // x32[2]
    if (precision == 3) b += 2LL * ld * 4 * n_wp * 2;                     // xb16[2]
    // the accumulators, coefficient rows and term ring of dh_propagate_steps'
    // multi-step series (allocated on first use)
    b += ((long long)MULTI_MAX * (ld * 2 * n_wp + KMAX_ALLOC + 2) + PANEL * ld * 2 * n_wp) * 16;
    return b;
}

extern "C" int dh_system_create(dh_context* ctx, int n_atoms, const int* first_orb, const int* atom_lmax, int n_orb,
                                const int* orb_n, const int* orb_l, const int* orb_m, const int* orb_nzeta,
                                const double* orb_zeta, const double* orb_coef, const double* orb_ip,
                                const double* orb_kwh, const double* orb_vshift, const int* orb_frag, int n_frag,
                                int n_wp, dh_system** out) {
    CHECK_ARG(ctx && out, "dh_system_create: null context or output");
    CHECK_ARG(n_orb > 0 && n_wp > 0 && n_wp <= 64, "dh_system_create: need n_orb > 0 and 1 <= n_wp <= 64");
    CHECK_ARG(n_frag >= 1 && n_frag <= 16, "dh_system_create: 1 <= n_frag <= 16");
    CHECK_ARG(n_atoms >= 0, "dh_system_create: n_atoms < 0");
    if (orb_frag)
        for (int i = 0; i < n_orb; ++i)
            CHECK_ARG(orb_frag[i] >= 0 && orb_frag[i] < n_frag, "dh_system_create: fragment index out of range");
    DH_HIP(hipSetDevice(ctx->device));
    std::vector<int> sh_n, sh_l, sh_nz, sh_orb0, at_ns, at_s0;
    std::vector<double> sh_z, sh_c;
    if (n_atoms > 0) {
        CHECK_ARG(first_orb && atom_lmax && orb_n && orb_l && orb_m && orb_nzeta && orb_zeta && orb_coef && orb_ip &&
                      orb_kwh && orb_vshift,
                  "dh_system_create: basis arrays missing");
        CHECK_ARG(first_orb[0] == 0 && first_orb[n_atoms] == n_orb, "dh_system_create: first_orb must span 0..n_orb");
        for (int a = 0; a < n_atoms; ++a) {
            at_s0.push_back((int)sh_n.size());
            int k = first_orb[a], ns = 0;
            CHECK_ARG(first_orb[a + 1] >= k, "dh_system_create: first_orb not monotone");
            while (k < first_orb[a + 1]) {
                const int l = orb_l[k];
                CHECK_ARG(l >= 0 && l <= 2, "dh_system_create: only s, p, d shells (l <= 2) are supported");
                CHECK_ARG(orb_n[k] >= 1 && orb_n[k] <= 5 && orb_n[k] > l, "dh_system_create: need l < n <= 5");
                CHECK_ARG(orb_nzeta[k] >= 1 && orb_nzeta[k] <= 2, "dh_system_create: Nzeta must be 1 or 2");
                CHECK_ARG(k + 2 * l + 1 <= first_orb[a + 1], "dh_system_create: incomplete shell");
                for (int m = -l; m <= l; ++m)
                    CHECK_ARG(orb_l[k + m + l] == l && orb_m[k + m + l] == m && orb_n[k + m + l] == orb_n[k],
                              "dh_system_create: shell orbitals must be ordered m = -l..l (structure.f:211-223)");
                sh_n.push_back(orb_n[k]);
                sh_l.push_back(l);
                sh_nz.push_back(orb_nzeta[k]);
                sh_orb0.push_back(k);
                sh_z.push_back(orb_zeta[2 * k]);
                sh_z.push_back(orb_zeta[2 * k + 1]);
                sh_c.push_back(orb_coef[2 * k]);
                sh_c.push_back(orb_coef[2 * k + 1]);
                k += 2 * l + 1;
                ++ns;
            }
            CHECK_ARG(ns <= MAX_SHELLS, "dh_system_create: at most 3 shells per atom");
            at_ns.push_back(ns);
        }
    }
    dh_system* s = new dh_system();
    s->ctx = ctx;
    s->n_atoms = n_atoms;
    s->N = n_orb;
    s->ld = round_ld(n_orb);
    s->W = n_wp;
    s->n_frag = n_frag;
    const int N = n_orb, ld = s->ld, R = 2 * n_wp;
    int rc = DH_OK;
#define SALLOC(p, n)                                   \
    do {                                               \
        if ((rc = dalloc(&(p), (n))) != DH_OK) {       \
            dh_system_destroy(s);                      \
            return rc;                                 \
        }                                              \
    } while (0)
    // basis: one allocation, carved
    const int ns = (int)sh_n.size();
    const size_t bytes = (size_t)n_atoms * 3 * sizeof(int) + (size_t)ns * 4 * sizeof(int) +
                         (size_t)ns * 4 * sizeof(double) + (size_t)N * 3 * sizeof(double) + 64 * 8;
    char* bm = nullptr;
    SALLOC(bm, bytes);
    s->basis_mem = bm;
    size_t off = 0;
    auto carve = [&](size_t nbytes) {
        char* p = bm + off;
        off += (nbytes + 15) & ~size_t(15);
        return p;
    };
    int* d_at_ns = (int*)carve(sizeof(int) * std::max(n_atoms, 1));
    int* d_at_s0 = (int*)carve(sizeof(int) * std::max(n_atoms, 1));
    int* d_order = (int*)carve(sizeof(int) * std::max(n_atoms, 1));
    int* d_sh_n = (int*)carve(sizeof(int) * std::max(ns, 1));
    int* d_sh_l = (int*)carve(sizeof(int) * std::max(ns, 1));
    int* d_sh_nz = (int*)carve(sizeof(int) * std::max(ns, 1));
    int* d_sh_o = (int*)carve(sizeof(int) * std::max(ns, 1));
    double* d_sh_z = (double*)carve(sizeof(double) * 2 * std::max(ns, 1));
    double* d_sh_c = (double*)carve(sizeof(double) * 2 * std::max(ns, 1));
    double* d_ip = (double*)carve(sizeof(double) * N);
    double* d_kwh = (double*)carve(sizeof(double) * N);
    double* d_vs = (double*)carve(sizeof(double) * N);
    if (off > bytes) {
        dh_system_destroy(s);
        dh::set_error("dh_system_create: internal basis layout overflow");
        return DH_ERR_ARG;
    }
    hipStream_t st = ctx->stream;
    if (n_atoms > 0) {
        DH_HIP(hipMemcpyAsync(d_at_ns, at_ns.data(), sizeof(int) * n_atoms, hipMemcpyHostToDevice, st));
        DH_HIP(hipMemcpyAsync(d_at_s0, at_s0.data(), sizeof(int) * n_atoms, hipMemcpyHostToDevice, st));
        DH_HIP(hipMemcpyAsync(d_sh_n, sh_n.data(), sizeof(int) * ns, hipMemcpyHostToDevice, st));
        DH_HIP(hipMemcpyAsync(d_sh_l, sh_l.data(), sizeof(int) * ns, hipMemcpyHostToDevice, st));
        DH_HIP(hipMemcpyAsync(d_sh_nz, sh_nz.data(), sizeof(int) * ns, hipMemcpyHostToDevice, st));
        DH_HIP(hipMemcpyAsync(d_sh_o, sh_orb0.data(), sizeof(int) * ns, hipMemcpyHostToDevice, st));
        DH_HIP(hipMemcpyAsync(d_sh_z, sh_z.data(), sizeof(double) * 2 * ns, hipMemcpyHostToDevice, st));
        DH_HIP(hipMemcpyAsync(d_sh_c, sh_c.data(), sizeof(double) * 2 * ns, hipMemcpyHostToDevice, st));
        DH_HIP(hipMemcpyAsync(d_ip, orb_ip, sizeof(double) * N, hipMemcpyHostToDevice, st));
        DH_HIP(hipMemcpyAsync(d_kwh, orb_kwh, sizeof(double) * N, hipMemcpyHostToDevice, st));
        DH_HIP(hipMemcpyAsync(d_vs, orb_vshift, sizeof(double) * N, hipMemcpyHostToDevice, st));
        // atoms ordered by basis signature (shells' n, l, zetas): the S build maps
        // a wavefront onto pairs of one element pair, so its lanes take the same branches
        std::vector<int> order(n_atoms);
        for (int a = 0; a < n_atoms; ++a) order[a] = a;
        auto key = [&](int a) {
            std::vector<double> k;
            for (int q = 0; q < at_ns[a]; ++q) {
                const int sh = at_s0[a] + q;
                k.insert(k.end(), {(double)sh_n[sh], (double)sh_l[sh], sh_z[2 * sh], sh_z[2 * sh + 1]});
            }
            return k;
        };
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return key(x) < key(y); });
        DH_HIP(hipMemcpyAsync(d_order, order.data(), sizeof(int) * n_atoms, hipMemcpyHostToDevice, st));
        DH_HIP(hipStreamSynchronize(st));  // the host vectors above go out of scope
        SALLOC(s->coords, (size_t)n_atoms * 3);
    }
    s->bd = BasisDev{n_atoms, s->coords, d_at_ns, d_at_s0, d_sh_n, d_sh_l, d_sh_nz, d_sh_o, d_sh_z, d_sh_c,
                     d_ip, d_kwh, d_vs, d_order};
    SALLOC(s->frag, N);
    {
        std::vector<int> fr(N, 0);
        if (orb_frag)
            for (int i = 0; i < N; ++i) fr[i] = orb_frag[i];
        DH_HIP(hipMemcpyAsync(s->frag, fr.data(), sizeof(int) * N, hipMemcpyHostToDevice, st));
        DH_HIP(hipStreamSynchronize(st));
    }
    const size_t mat = (size_t)ld * ld;
    SALLOC(s->Q1, mat);
    SALLOC(s->Q2, mat);
    SALLOC(s->Hth, mat);
    SALLOC(s->phi, (size_t)ld * R);
    SALLOC(s->x0, (size_t)ld * R);
    SALLOC(s->x1, (size_t)ld * R);
    SALLOC(s->acc, (size_t)ld * R);
    SALLOC(s->out, (size_t)ld * 3 * n_wp);
    SALLOC(s->hk, (size_t)ld * n_wp);
    SALLOC(s->stage, (size_t)ld * 2);
    // graphs save launch overhead on short terms; for large N a re-capture after
    // every change of the spectral bounds costs more than it saves
    s->use_graph = N <= 2048;
    s->split_max = matvec_split_for(N, 1);
    SALLOC(s->partial, (size_t)partial_cplx_for(N, n_wp));
    SALLOC(s->counters, (size_t)ld);
    SALLOC(s->lv, (size_t)ld);
    SALLOC(s->lvp, (size_t)ld);
    SALLOC(s->lw, (size_t)ld);
    SALLOC(s->lal, LANCZOS_MAX + 2);
    SALLOC(s->lbe, LANCZOS_MAX + 2);
    SALLOC(s->J, KMAX_ALLOC + 2);
    SALLOC(s->coef, KMAX_ALLOC + 2);
    SALLOC(s->kout, 4);
    SALLOC(s->pops, (size_t)n_wp * (n_frag + 1));
    SALLOC(s->erg, (size_t)n_wp * 2);
    SALLOC(s->norms, (size_t)n_wp * 2);
    SALLOC(s->norms0, (size_t)R);
    SALLOC(s->norms1, (size_t)R);
    DH_HIP(hipEventCreate(&s->ev0));
    DH_HIP(hipEventCreate(&s->ev1));
    launch_zero_u32(s->counters, ld, st);
    launch_zero_cplx(s->out, (long)ld * 3 * n_wp, st);
    launch_zero_cplx(s->phi, (long)ld * R, st);
    launch_zero_cplx(s->hk, (long)ld * n_wp, st);
    launch_zero_cplx(s->lw, (long)ld, st);
    DH_HIP(hipGetLastError());
    DH_HIP(hipStreamSynchronize(st));
    *out = s;
    return DH_OK;
}

extern "C" int dh_set_coords(dh_system* s, const double* coords) {
    CHECK_ARG(s && coords, "dh_set_coords: null argument");
    CHECK_ARG(s->n_atoms > 0, "dh_set_coords: operator-only system has no atoms");
    DH_HIP(hipSetDevice(s->ctx->device));
    DH_HIP(hipMemcpyAsync(s->coords, coords, sizeof(double) * 3 * s->n_atoms, hipMemcpyHostToDevice, s->ctx->stream));
    DH_HIP(hipStreamSynchronize(s->ctx->stream));
    s->have_coords = true;
    return DH_OK;
}

// phi <- (L^-1 bra, L^T ket) from the AO state held in out[:, 0:W] and out[:, 2W:3W]
static void convert_ao_to_phi(dh_system* s) {
    hipStream_t st = s->ctx->stream;
    const int W = s->W;
    MatvecArgs a{};
    a.ld = s->ld;
    a.N = s->N;
    a.partial = s->partial;
    a.counters = s->counters;
    a.epi = EPI_PLAIN;
    a.R = W;
    a.split = matvec_split_for(s->N, W);
    // phi_bra = L^-1 bra : Q2 = sym(L^-1), j <= i
    a.M = s->Q2;
    a.tri = TRI_UPPER;
    a.x = s->out; a.x_stride = 3 * W; a.x_off = 0;
    a.out = s->phi; a.out_stride = 2 * W; a.out_off = 0;
    launch_matvec(a, st);
    // phi_ket = L^T ket : Q1 = sym(L), j >= i
    a.M = s->Q1;
    a.tri = TRI_LOWER;
    a.x = s->out; a.x_stride = 3 * W; a.x_off = 2 * W;
    a.out = s->phi; a.out_stride = 2 * W; a.out_off = W;
    launch_matvec(a, st);
}

// Buffers one operator build writes and the stream it runs on: the system's own
// operators with the context workspace (a build in place), or the shadow set
// that a build overlapped with propagation fills on the side stream.
struct OpsTarget {
    hipStream_t st;
    double *S, *h, *W, *Dinv;
    int* info;
    double *Q1, *Q2, *Hth;
    float* H32;
    unsigned short* Hlo;
    cplx *lv, *lvp, *lw, *partial;
    double *lal, *lbe;
    unsigned* counters;
    double* bounds;  // 6 values
};

static OpsTarget main_target(dh_system* s) {
    dh_context* c = s->ctx;
    return OpsTarget{c->stream, c->S, c->h, c->W, c->Dinv, c->d_info, s->Q1, s->Q2, s->Hth, s->H32, s->Hlo,
                     s->lv, s->lvp, s->lw, s->partial, s->lal, s->lbe, s->counters, s->bounds_h};
}

// Lanczos spectral bounds of t.Hth into t.bounds (host wait on t.st)
// With al_host/be_host (pinned), the coefficients are only copied there on
// t.st and the host part (lanczos_bounds_host) is left to the caller, after
// the stream has passed them; *m_out gets their count.
static int bounds_into(dh_system* s, const OpsTarget& t, double* al_host = nullptr, double* be_host = nullptr,
                       int* m_out = nullptr) {
    hipStream_t st = t.st;
    const int m = std::min(s->lanczos_steps, std::min(LANCZOS_MAX, s->N));
    launch_lanczos_start(s->N, t.lv, t.lvp, 0x9e3779b9u, st);
    MatvecArgs a{};
    a.M = t.Hth;
    a.ld = s->ld;
    a.N = s->N;
    a.R = 1;
    a.partial = t.partial;
    a.counters = t.counters;
    a.split = matvec_split_for(s->N, 1);
    a.tri = TRI_FULL;
    a.epi = EPI_PLAIN;
    a.x = t.lv; a.x_stride = 1; a.x_off = 0;
    a.out = t.lw; a.out_stride = 1; a.out_off = 0;
    for (int j = 0; j < m; ++j) {
        launch_matvec(a, st);
        launch_lanczos_step(s->N, t.lv, t.lvp, t.lw, t.lal, t.lbe, j, st);
    }
    DH_HIP(hipGetLastError());
    if (al_host) {
        DH_HIP(hipMemcpyAsync(al_host, t.lal, sizeof(double) * m, hipMemcpyDeviceToHost, st));
        DH_HIP(hipMemcpyAsync(be_host, t.lbe, sizeof(double) * (m + 1), hipMemcpyDeviceToHost, st));
        *m_out = m;
        return DH_OK;
    }
    double al[LANCZOS_MAX], be[LANCZOS_MAX + 1];
    DH_HIP(hipMemcpyAsync(al, t.lal, sizeof(double) * m, hipMemcpyDeviceToHost, st));
    DH_HIP(hipMemcpyAsync(be, t.lbe, sizeof(double) * (m + 1), hipMemcpyDeviceToHost, st));
    DH_HIP(hipStreamSynchronize(st));
    lanczos_bounds_host(al, be, m, s->rel_margin, t.bounds);
    return DH_OK;
}

// the bounds of a finished build become the system's
static void take_bounds(dh_system* s, const double* b6) {
    for (int i = 0; i < 6; ++i) s->bounds_h[i] = b6[i];
    s->emin = s->bounds_h[0];
    s->emax = s->bounds_h[1];
    s->widen = 1.0;
    s->coef_dt = -1.0;  // force new coefficients
}

static int compute_bounds(dh_system* s) {
    OpsTarget t = main_target(s);
    int rc = bounds_into(s, t);
    if (rc) return rc;
    take_bounds(s, s->bounds_h);
    return DH_OK;
}

// Project autotuner: time every matvec kernel variant on this system's real
// operands for each shape the propagation uses, keep the fastest.  Variants
// differ only in tile shape and reduction tree (all FP64, all deterministic).
static int autotune_shape(dh_system* s, const MatvecArgs& base) {
    hipStream_t st = s->ctx->stream;
    int best = -1;
    float best_ms = 1e30f;
    const int nv = matvec_variant_count();
    for (int v = 0; v < nv; ++v) {
        for (int w = 0; w < 2; ++w) launch_matvec_variant(base, v, st);
        DH_HIP(hipEventRecord(s->ev0, st));
        const int reps = 6;
        for (int r = 0; r < reps; ++r) launch_matvec_variant(base, v, st);
        DH_HIP(hipEventRecord(s->ev1, st));
        DH_HIP(hipEventSynchronize(s->ev1));
        float ms = 0.f;
        DH_HIP(hipEventElapsedTime(&ms, s->ev0, s->ev1));
        if (ms < best_ms) {
            best_ms = ms;
            best = v;
        }
    }
    matvec_set_tuned(base.N, base.R, base.tri, best);
    return DH_OK;
}

static int autotune_system(dh_system* s) {
    const int W = s->W, R = 2 * W;
    MatvecArgs a{};
    a.ld = s->ld;
    a.N = s->N;
    a.partial = s->partial;
    a.counters = s->counters;
    a.epi = EPI_PLAIN;
    a.x_off = 0;
    a.out_off = 0;
    struct Shape {
        const double* M;
        int tri, R;
        cplx *x, *out;
        int xs, os;
    } shapes[] = {
        {s->Hth, TRI_FULL, R, s->phi, s->x0, R, R},   // Chebyshev terms
        {s->Hth, TRI_FULL, W, s->phi, s->x1, R, W},   // <H> matvec
        {s->Q1, TRI_UPPER, W, s->phi, s->x1, R, W},   // L phi_bra
        {s->Q2, TRI_LOWER, R, s->phi, s->x0, R, R},   // L^-T phi
        {s->Q1, TRI_LOWER, W, s->phi, s->x1, R, W},   // L^T ket
        {s->Q2, TRI_UPPER, W, s->phi, s->x1, R, W},   // L^-1 bra
        {s->Hth, TRI_FULL, 1, s->lv, s->lw, 1, 1},    // Lanczos
    };
    for (auto& sh : shapes) {
        a.M = sh.M;
        a.tri = sh.tri;
        a.R = sh.R;
        a.x = sh.x;
        a.out = sh.out;
        a.x_stride = sh.xs;
        a.out_stride = sh.os;
        if (matvec_is_tuned(s->N, sh.R, sh.tri)) continue;
        a.split = matvec_split_for(s->N, sh.R);
        int rc = autotune_shape(s, a);
        if (rc) return rc;
    }
    s->tuned = true;
    return DH_OK;
}

// H' = S^-1 h = L^-T W into the context's Hp, from W = L^-1 h.  linv holds
// L^-1 with a zero upper triangle: the product reads whole diagonal tiles.
static void launch_hprime(dh_system* s, const double* linv) {
    dh_context* c = s->ctx;
    const int ld = s->ld;
    Gemm g{};
    g.M = ld; g.N = ld; g.K = ld;
    g.A = linv; g.lda = ld; g.transA = 1;
    g.B = c->W; g.ldb = ld;
    g.C = c->Hp; g.ldc = ld;
    g.alpha = 1.0; g.beta = 0.0; g.kmode = K_FROM_ROWTILE; g.batch = 1;
    launch_gemm(g, c->stream);
    s->hprime_ready = true;
}

// S = L L^T, L^-1, W = L^-1 h, H_orth = W L^-T from t.S and t.h, all on t.st
// (host wait for the Cholesky status); with hprime (in-place builds only) also
// H' = L^-T W into the context's Hp.  Q1 and Q2 come back mirrored.  With
// info_host (pinned) the status is only copied there on t.st and the host
// does not wait: the caller checks it once the stream has passed it (the
// products after a failed factorisation are then discarded).
static int factor_into(dh_system* s, const OpsTarget& t, bool hprime, int* info_host = nullptr) {
    hipStream_t st = t.st;
    const int ld = s->ld;
    launch_copy_matrix(t.S, t.Q1, (long)ld * ld, st);
    DH_HIP(hipMemsetAsync(t.info, 0, sizeof(int), st));
    launch_cholesky(t.Q1, ld, t.Dinv, t.info, st);
    if (info_host) {
        DH_HIP(hipMemcpyAsync(info_host, t.info, sizeof(int), hipMemcpyDeviceToHost, st));
    } else {
        int info = 0;
        DH_HIP(hipMemcpyAsync(&info, t.info, sizeof(int), hipMemcpyDeviceToHost, st));
        DH_HIP(hipStreamSynchronize(st));
        if (info != 0) {
            dh::set_error("dh_orthogonalize: S is not positive definite (Cholesky failed at column %d)", info);
            return DH_ERR_NOT_SPD;
        }
    }
    launch_trtri_lower(t.Q1, t.Q2, t.Dinv, ld, t.W, st);  // W is free until W = L^-1 h
    {   // W = L^-1 h
        Gemm g{};
        g.M = ld; g.N = ld; g.K = ld;
        g.A = t.Q2; g.lda = ld;
        g.B = t.h; g.ldb = ld;
        g.C = t.W; g.ldc = ld;
        g.alpha = 1.0; g.beta = 0.0; g.kmode = K_BELOW_ROWTILE; g.batch = 1;
        launch_gemm(g, st);
    }
    {   // H_orth = W L^-T (lower tiles), then mirror
        Gemm g{};
        g.M = ld; g.N = ld; g.K = ld;
        g.A = t.W; g.lda = ld;
        g.B = t.Q2; g.ldb = ld; g.transB = 1;
        g.C = t.Hth; g.ldc = ld;
        g.alpha = 1.0; g.beta = 0.0; g.kmode = K_BELOW_COLTILE; g.lower_only = 1; g.batch = 1;
        launch_gemm(g, st);
        launch_symmetrize_from_lower(t.Hth, ld, st);
    }
    if (hprime) launch_hprime(s, t.Q2);  // Q2 is not mirrored yet
    launch_symmetrize_from_lower(t.Q1, ld, st);
    launch_symmetrize_from_lower(t.Q2, ld, st);
    DH_HIP(hipGetLastError());
    return DH_OK;
}

// the FP32 copies of t.Hth that the precision modes read
static void f32_copies_into(dh_system* s, const OpsTarget& t) {
    const int ld = s->ld;
    if (s->precision == 3) launch_split_f32_bf16(t.Hth, t.H32, t.Hlo, (long)ld * ld, t.st);
    else if (s->precision != 0) launch_to_f32(t.Hth, t.H32, (long)ld * ld, t.st);
}

extern "C" int dh_orthogonalize(dh_system* s, int want_hprime) {
    CHECK_ARG(s, "dh_orthogonalize: null system");
    dh_context* c = s->ctx;
    if (c->ws_owner != s || c->ws_ld < s->ld) {
        dh::set_error("dh_orthogonalize: S and h of this system are not loaded (call dh_build_operators or dh_load_operators)");
        return DH_ERR_STATE;
    }
    DH_HIP(hipSetDevice(c->device));
    OpsTarget t = main_target(s);
    int rc = factor_into(s, t, want_hprime != 0);
    if (rc) {
        s->ops_ready = false;
        return rc;
    }
    s->hprime_ready = want_hprime != 0;
    s->w_ready = true;
    if (!s->tuned) {
        int trc = autotune_system(s);
        if (trc) return trc;
    }
    f32_copies_into(s, t);
    s->h32_ready = s->precision != 0;
    rc = compute_bounds(s);
    if (rc) return rc;
    s->ops_ready = true;
    s->phi_dirty = true;
    return DH_OK;
}

extern "C" int dh_build_operators(dh_system* s, double cutoff, double s_drop) {
    CHECK_ARG(s, "dh_build_operators: null system");
    s->stats[7] += 1;
    CHECK_ARG(s->n_atoms > 0 && s->have_coords, "dh_build_operators: coordinates not set");
    DH_HIP(hipSetDevice(s->ctx->device));
    int rc = ensure_ws(s->ctx, s->ld);
    if (rc) return rc;
    s->ctx->ws_owner = s;
    if (launch_build_SH(s->bd, s->N, s->ld, s->ctx->S, s->ctx->h, cutoff, s_drop, s->ctx->stream)) {
        dh::set_error("dh_build_operators: kernel launch failed: %s", hipGetErrorString(hipGetLastError()));
        return DH_ERR_HIP;
    }
    return dh_orthogonalize(s, 0);
}

// Host <-> device transfers as contiguous copies: a strided host<->device copy
// with small rows goes row by row through pageable staging, which is slow at
// large N.  Complex columns pass through the system's staging columns and a
// strided copy kernel; matrices are copied whole when the pitches agree and
// column by column otherwise.
static int put_column(dh_system* s, cplx* dst, long dst_stride, const double* host, int slot) {
    hipStream_t st = s->ctx->stream;
    cplx* stg = s->stage + (size_t)slot * s->ld;
    DH_HIP(hipMemcpyAsync(stg, host, sizeof(cplx) * s->N, hipMemcpyHostToDevice, st));
    launch_strided_copy(dst, dst_stride, stg, 1, s->N, st);
    return DH_OK;
}

static int get_column(dh_system* s, double* host, const cplx* src, long src_stride, int slot) {
    hipStream_t st = s->ctx->stream;
    cplx* stg = s->stage + (size_t)slot * s->ld;
    launch_strided_copy(stg, 1, src, src_stride, s->N, st);
    DH_HIP(hipMemcpyAsync(host, stg, sizeof(cplx) * s->N, hipMemcpyDeviceToHost, st));
    return DH_OK;
}

// rows x cols column-major block between pitches ldd and lds
static int copy_block(double* dst, long ldd, const double* src, long lds, int rows, int cols, hipMemcpyKind kind,
                      hipStream_t st) {
    DH_HIP(hipMemcpy2DAsync(dst, sizeof(double) * ldd, src, sizeof(double) * lds, sizeof(double) * rows, cols, kind,
                            st));
    return DH_OK;
}

// n x n matrix
static int copy_matrix(double* dst, long ldd, const double* src, long lds, int n, hipMemcpyKind kind,
                       hipStream_t st) {
    if (ldd == lds) {
        DH_HIP(hipMemcpyAsync(dst, src, sizeof(double) * (size_t)ldd * n, kind, st));
        return DH_OK;
    }
    for (int j = 0; j < n; ++j)
        DH_HIP(hipMemcpyAsync(dst + (size_t)ldd * j, src + (size_t)lds * j, sizeof(double) * n, kind, st));
    return DH_OK;
}

extern "C" int dh_load_operators(dh_system* s, const double* S, const double* h, int ldh) {
    CHECK_ARG(s && S && h, "dh_load_operators: null argument");
    CHECK_ARG(ldh >= s->N, "dh_load_operators: ld < N");
    DH_HIP(hipSetDevice(s->ctx->device));
    int rc = ensure_ws(s->ctx, s->ld);
    if (rc) return rc;
    hipStream_t st = s->ctx->stream;
    s->ctx->ws_owner = s;
    launch_identity_pad(s->ctx->S, s->ctx->h, s->N, s->ld, st);
    // N x N only: rows N..ld-1 keep the identity padding (a whole-column copy
    // would bring in the caller's padding rows)
    const bool whole = ldh == s->N && s->ld == s->N;
    for (int j = 0; j < (whole ? 1 : s->N); ++j) {
        const size_t cnt = whole ? (size_t)s->N * s->N : (size_t)s->N;
        DH_HIP(hipMemcpyAsync(s->ctx->S + (size_t)s->ld * j, S + (size_t)ldh * j, sizeof(double) * cnt,
                              hipMemcpyHostToDevice, st));
        DH_HIP(hipMemcpyAsync(s->ctx->h + (size_t)s->ld * j, h + (size_t)ldh * j, sizeof(double) * cnt,
                              hipMemcpyHostToDevice, st));
    }
    return DH_OK;
}

extern "C" int dh_get_matrix(dh_system* s, int which, double* out, int ldo) {
    CHECK_ARG(s && out && ldo >= s->N, "dh_get_matrix: bad argument");
    DH_HIP(hipSetDevice(s->ctx->device));
    const double* src = nullptr;
    const bool ws_ok = s->ctx->ws_owner == s && s->ctx->ws_ld >= s->ld;
    switch (which) {
        case DH_MAT_S: src = ws_ok ? s->ctx->S : nullptr; break;
        case DH_MAT_H: src = ws_ok ? s->ctx->h : nullptr; break;
        case DH_MAT_L: src = s->Q1; break;
        case DH_MAT_LINV: src = s->Q2; break;
        case DH_MAT_HORTH: src = s->Hth; break;
        case DH_MAT_HPRIME: src = (s->hprime_ready && ws_ok) ? s->ctx->Hp : nullptr; break;
        default: break;
    }
    if (!src) {
        dh::set_error("dh_get_matrix: matrix %d not available", which);
        return DH_ERR_STATE;
    }
    int rc = copy_matrix(out, ldo, src, s->ld, s->N, hipMemcpyDeviceToHost, s->ctx->stream);
    if (rc) return rc;
    DH_HIP(hipStreamSynchronize(s->ctx->stream));
    return DH_OK;
}

extern "C" int dh_get_bounds(dh_system* s, double* out6) {
    CHECK_ARG(s && out6, "dh_get_bounds: null argument");
    CHECK_ARG(s->ops_ready, "dh_get_bounds: operators not built");
    for (int i = 0; i < 6; ++i) out6[i] = s->bounds_h[i];
    out6[0] = s->emin;
    out6[1] = s->emax;
    return DH_OK;
}

extern "C" int dh_set_lanczos(dh_system* s, int steps, double rel_margin) {
    CHECK_ARG(s && steps >= 4 && steps <= LANCZOS_MAX && rel_margin >= 0.0, "dh_set_lanczos: bad argument");
    s->lanczos_steps = steps;
    s->rel_margin = rel_margin;
    return DH_OK;
}

extern "C" int dh_set_wavepacket(dh_system* s, int wp, const double* bra, const double* ket) {
    CHECK_ARG(s && bra && ket && wp >= 0 && wp < s->W, "dh_set_wavepacket: bad argument");
    DH_HIP(hipSetDevice(s->ctx->device));
    const int W = s->W;
    hipStream_t st = s->ctx->stream;
    int rc = put_column(s, s->out + wp, 3 * W, bra, 0);
    if (rc) return rc;
    rc = put_column(s, s->out + 2 * W + wp, 3 * W, ket, 1);
    if (rc) return rc;
    DH_HIP(hipGetLastError());
    DH_HIP(hipStreamSynchronize(st));
    s->phi_dirty = true;
    s->state_set = true;
    return DH_OK;
}

extern "C" int dh_get_wavepacket(dh_system* s, int wp, double* bra, double* ket) {
    CHECK_ARG(s && wp >= 0 && wp < s->W, "dh_get_wavepacket: bad argument");
    DH_HIP(hipSetDevice(s->ctx->device));
    const int W = s->W;
    hipStream_t st = s->ctx->stream;
    int rc = DH_OK;
    if (bra && (rc = get_column(s, bra, s->out + wp, 3 * W, 0))) return rc;
    if (ket && (rc = get_column(s, ket, s->out + 2 * W + wp, 3 * W, 1))) return rc;
    DH_HIP(hipGetLastError());
    DH_HIP(hipStreamSynchronize(st));
    return DH_OK;
}

static int ensure_coefficients(dh_system* s, double dt) {
    const double c = 0.5 * (s->emax + s->emin);
    const double r = 0.5 * (s->emax - s->emin) * s->widen;
    if (dt == s->coef_dt && c == s->coef_c && r == s->coef_r) return DH_OK;
    hipStream_t st = s->ctx->stream;
    // on the host (a few microseconds; the device recurrence is one thread of
    // dependent memory accesses, about a millisecond), then one copy on the
    // stream: the host neither launches nor waits for a kernel here
    if (!s->coef_h) {
        DH_HIP(hipHostMalloc((void**)&s->coef_h, sizeof(cplx) * (KMAX_ALLOC + 2), hipHostMallocDefault));
        DH_HIP(hipEventCreateWithFlags(&s->coef_ev, hipEventDisableTiming));
    }
    if (s->coef_pending) DH_HIP(hipEventSynchronize(s->coef_ev));  // coef_h is free again
    const int k = bessel_coef_host(r * dt, c * dt, KMAX_ALLOC, BESSEL_TOL, s->coef_h);
    if (k <= 0 && r * dt > 1e-12) {
        dh::set_error("Chebyshev coefficients: no terms for r*dt = %g", r * dt);
        return DH_ERR_STATE;
    }
    DH_HIP(hipMemcpyAsync(s->coef, s->coef_h, sizeof(cplx) * (k + 2), hipMemcpyHostToDevice, st));
    DH_HIP(hipEventRecord(s->coef_ev, st));
    s->coef_pending = true;
    s->K = k;
    s->coef_dt = dt;
    s->coef_c = c;
    s->coef_r = r;
    return DH_OK;
}

// Operands of one Chebyshev evaluation acc <- sum_k a_k T_k(Hh) phi, Hh = (H - c)/r,
// with the device coefficients a_0..a_K already in coef.  phi is read only;
// the recurrence alternates between x0 and x1.
struct SeriesOps {
    const double* H;
    const float* H32;          // precision modes 1, 2 and 3
    const unsigned short* Hlo;  // precision mode 3
    int ld, N, R;
    const cplx* phi;
    cplx *x0, *x1, *acc, *partial;
    unsigned* counters;
    const cplx* coef;
    float* x32[2];              // precision modes 1 and 3
    unsigned short* xb16[2];    // precision mode 3
    LibHandle* lib;             // precision modes 1 and 3
    int prec, K;
    double c, r;
    // a multi-step series (m > 0): accumulator q at accm + q * accm_stride takes
    // the coefficients at coefm + q * coefm_stride, terms 0..kq[q] (device
    // table; Kq on the host); the terms go to the PANEL slots of ring
    // (ring_stride apart) and reach the accumulators term by term or in panel
    // passes (series_core).  K is then kq[m - 1]; acc, coef, x0 and x1 are not
    // used.
    int m = 0;
    cplx* accm = nullptr;
    long accm_stride = 0;
    const cplx* coefm = nullptr;
    int coefm_stride = 0;
    const int* kq = nullptr;
    const int* Kq = nullptr;  // the same table on the host
    cplx* ring = nullptr;
    long ring_stride = 0;
};


static int series_core(const SeriesOps& o, hipStream_t st) {
    const int R = o.R;
    const double c = o.c, r = o.r;
    MatvecArgs a{};
    a.M = o.H;
    a.ld = o.ld;
    a.N = o.N;
    a.R = R;
    a.partial = o.partial;
    a.counters = o.counters;
    a.split = matvec_split_for(o.N, R);
    a.tri = TRI_FULL;
    a.x_stride = R; a.x_off = 0;
    a.out_stride = R; a.out_off = 0;
    const bool panels = o.m > 0;
    a.acc = panels ? nullptr : o.acc;
    a.coef = panels ? nullptr : o.coef;
    a.M32 = o.H32;
    // a multi-step series: a term whose phase 2 is the symmetric block
    // reduction adds T_k to the accumulators itself (small R: the update rides
    // on loads the epilogue makes anyway); after any other term kind the
    // accumulators take the terms in panel passes (one read of each T_k for
    // all accumulators).  The first term decides; every term of a series runs
    // the same kernel kind.
    if (panels) {
        a.accm = o.accm;
        a.accm_stride = o.accm_stride;
        a.coefm = o.coefm;
        a.coefm_stride = o.coefm_stride;
        a.m_lo = 0;
        a.m_hi = o.m;
    }
    const PanelArgs pa{o.accm, o.accm_stride, o.coefm, o.coefm_stride, o.ring, o.ring_stride, o.phi, o.kq, o.m,
                       (long)o.N * R};
    auto slot = [&](int k) { return o.ring + (long)((k - 1) % PANEL) * o.ring_stride; };
    int k_taken = -1;   // the last term the accumulators hold
    int per_term = -1;  // -1: not known yet; 1: the terms accumulate; 0: panel passes
    auto accumulated = [&](int k) -> int {
        if (!panels) return DH_OK;
        const int done = take_multi_done() ? 1 : 0;
        if (per_term < 0) per_term = done;
        if (done != per_term) {
            dh::set_error("Chebyshev series: term %d changed kernel kind within a multi-step series", k);
            return DH_ERR_STATE;
        }
        if (per_term) {
            k_taken = k;
        } else if (k % PANEL == 0 || k == o.K) {
            launch_multi_panel(pa, k_taken + 1, k, st);
            k_taken = k;
        }
        return DH_OK;
    };
    const int prec = o.prec;
    // modes 1 and 3: the FP32 products read FP32 copies of their input (and mode
    // 3's residual products bf16 copies); the step's state is converted once,
    // then each term's epilogue writes the next one's
    const bool f32x = prec == 1 || prec == 3;
    const long nx = (long)o.ld * 2 * R;
    if (f32x) launch_to_f32(reinterpret_cast<const double*>(o.phi), o.x32[0], nx, st);
    if (prec == 3) launch_to_bf16(o.x32[0], o.xb16[0], nx, st);
    float* Phi = reinterpret_cast<float*>(o.partial);
    float* Plo = Phi + (long)LIB_SLICES * 2 * R * o.ld;
    auto term = [&](MatvecArgs args) -> int {
        if (f32x) {
            const int in = (args.k - 1) & 1;
            args.x32 = o.x32[in];
            args.out32 = o.x32[args.k & 1];
            if (prec == 3) args.outb16 = o.xb16[args.k & 1];
            const int KS = o.ld / LIB_SLICES;
            int rc = lib_gemm_slices_f32(o.lib, o.ld, 2 * R, KS, LIB_SLICES, o.H32, o.ld, o.x32[in], 2 * R, Phi,
                                         o.ld);
            if (!rc && prec == 3)
                rc = lib_gemm_slices_bf16(o.lib, o.ld, 2 * R, KS, LIB_SLICES, o.Hlo, o.ld, o.xb16[in], 2 * R, Plo,
                                          o.ld);
            launch_lib_reduce(args, Phi, prec == 3 ? Plo : nullptr, LIB_SLICES, st);
            return rc;
        } else if (prec == 2) launch_term_f32h(args, st);
        else launch_matvec(args, st);
        return DH_OK;
    };
    // k = 1 (for K = 0 the device table holds a_1 = 0 and acc = a_0 phi)
    a.epi = EPI_CHEB_FIRST;
    a.x = o.phi;
    a.out = panels ? slot(1) : o.x1;
    a.alpha = 1.0 / r;
    a.beta = -c / r;
    a.k = 1;
    take_multi_done();
    if (int rc = term(a)) return rc;
    if (int rc = accumulated(1)) return rc;
    const cplx* prev2 = o.phi;
    cplx* prev = a.out;
    for (int k = 2; k <= o.K; ++k) {
        cplx* outb = panels ? slot(k) : (k == 2) ? o.x0 : const_cast<cplx*>(prev2);
        // a term kernel that accumulates takes the terms two at a time: odd k
        // adds T_{k-1} (its input) and T_k, even k nothing, except a last
        // even k, which adds T_k alone; accumulators whose series has ended
        // before the first term added take no more
        a.m_pair = (k & 1) ? 2 : (k == o.K ? 1 : 0);
        if (panels && a.m_pair > 0)
            while (a.m_lo < o.m - 1 && o.Kq[a.m_lo] < k - (a.m_pair == 2 ? 1 : 0)) ++a.m_lo;
        a.epi = EPI_CHEB;
        a.x = prev;
        a.x2 = prev2;
        a.out = outb;
        a.alpha = 2.0 / r;
        a.beta = -2.0 * c / r;
        a.k = k;
        if (int rc = term(a)) return rc;
        if (int rc = accumulated(k)) return rc;
        prev2 = prev;
        prev = outb;
    }
    return DH_OK;
}

// One Chebyshev evaluation phi <- sum_k a_k T_k(Hh) phi, result in s->acc.
static int series_launches(dh_system* s) {
    SeriesOps o{};
    o.H = s->Hth;
    o.H32 = s->H32;
    o.Hlo = s->Hlo;
    o.ld = s->ld;
    o.N = s->N;
    o.R = 2 * s->W;
    o.phi = s->phi;
    o.x0 = s->x0;
    o.x1 = s->x1;
    o.acc = s->acc;
    o.partial = s->partial;
    o.counters = s->counters;
    o.coef = s->coef;
    o.x32[0] = s->x32[0];
    o.x32[1] = s->x32[1];
    o.xb16[0] = s->xb16[0];
    o.xb16[1] = s->xb16[1];
    o.lib = s->ctx->lib;
    o.prec = s->precision;
    o.K = s->K;
    o.c = s->coef_c;
    o.r = s->coef_r;
    if (int rc = series_core(o, s->ctx->stream)) return rc;
    DH_HIP(hipGetLastError());
    return DH_OK;
}

// The term series as one hipGraph launch: captured once per (K, bounds, buffer
// assignment, precision, term kernel) and replayed on later steps.  The series
// is kernels only (no memset nodes, no host synchronisation).
static int chebyshev_series(dh_system* s) {
    // the FP32-product modes call rocBLAS, which stays out of stream capture
    if (!s->use_graph || fp32_products(s)) return series_launches(s);
    hipStream_t st = s->ctx->stream;
    const int variant = matvec_variant_for(s->N, 2 * s->W, TRI_FULL);
    for (auto& g : s->graphs)
        if (g.exec && g.K == s->K && g.phi == s->phi && g.H == s->Hth && g.c == s->coef_c && g.r == s->coef_r &&
            g.prec == s->precision && g.variant == variant) {
            DH_HIP(hipGraphLaunch(g.exec, st));
            return DH_OK;
        }
    auto& g = s->graphs[s->next_graph];
    s->next_graph ^= 1;
    if (g.exec) {
        DH_HIP(hipGraphExecDestroy(g.exec));
        g.exec = nullptr;
    }
    DH_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int rc = series_launches(s);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &graph);
    if (rc) {
        if (graph) hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) {
        dh::set_error("term series capture failed: %s", hipGetErrorString(e));
        return DH_ERR_HIP;
    }
    const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (ei != hipSuccess) {
        g.exec = nullptr;
        dh::set_error("term series graph instantiation failed: %s", hipGetErrorString(ei));
        return DH_ERR_HIP;
    }
    g.K = s->K;
    g.phi = s->phi;
    g.H = s->Hth;
    g.c = s->coef_c;
    g.r = s->coef_r;
    g.prec = s->precision;
    g.variant = variant;
    DH_HIP(hipGraphLaunch(g.exec, st));
    return DH_OK;
}

static void launch_outputs(dh_system* s) {
    hipStream_t st = s->ctx->stream;
    const int W = s->W;
    MatvecArgs a{};
    a.ld = s->ld;
    a.N = s->N;
    a.partial = s->partial;
    a.counters = s->counters;
    a.epi = EPI_PLAIN;
    // hk = H_orth phi_ket
    a.M = s->Hth; a.tri = TRI_FULL; a.R = W; a.split = matvec_split_for(s->N, W);
    a.x = s->phi; a.x_stride = 2 * W; a.x_off = W;
    a.out = s->hk; a.out_stride = W; a.out_off = 0;
    launch_matvec(a, st);
    // bra = L phi_bra : Q1 = sym(L), j <= i
    a.M = s->Q1; a.tri = TRI_UPPER; a.R = W;
    a.x = s->phi; a.x_stride = 2 * W; a.x_off = 0;
    a.out = s->out; a.out_stride = 3 * W; a.out_off = 0;
    launch_matvec(a, st);
    // [AO_bra | ket] = L^-T [phi_bra | phi_ket] : Q2 = sym(L^-1), j >= i
    a.M = s->Q2; a.tri = TRI_LOWER; a.R = 2 * W; a.split = matvec_split_for(s->N, 2 * W);
    a.x = s->phi; a.x_stride = 2 * W; a.x_off = 0;
    a.out = s->out; a.out_stride = 3 * W; a.out_off = W;
    launch_matvec(a, st);
    launch_observables(s->N, W, s->n_frag, s->frag, s->out, s->phi, s->hk, s->pops, s->erg, s->norms, st);
}

// One step of dt (eV^-1) with the norm check: the term series, the new state's
// norms and the outputs, then one host wait; a step whose norm drifts is
// redone with wider bounds (at most 4 attempts).  phi holds the current state.
static int propagate_checked(dh_system* s, double dt, int* n_terms) {
    hipStream_t st = s->ctx->stream;
    const int R = 2 * s->W;
    launch_column_norms(s->N, R, s->phi, s->norms0, st);
    std::vector<double> n0(R), n1(R);
    // FP64 norm drift of one step that signals bounds not enclosing the
    // spectrum; the mixed mode's FP32 products drift ~1e-7 per step by themselves
    // (mode 2 propagates with the exactly symmetric FP32-rounded H: FP64 drift)
    const double drift_tol = fp32_products(s) ? 1e-4 : 1e-9;
    for (int attempt = 0; attempt < 4; ++attempt) {
        int rc = ensure_coefficients(s, dt);
        if (rc) return rc;
        if (s->timing) DH_HIP(hipEventRecord(s->ev0, st));
        rc = chebyshev_series(s);
        if (rc) return rc;
        if (s->timing) DH_HIP(hipEventRecord(s->ev1, st));
        std::swap(s->phi, s->acc);  // phi <- new state; acc keeps the old one
        launch_column_norms(s->N, R, s->phi, s->norms1, st);
        // the outputs go out with the norms, so one host wait covers the step;
        // an attempt that fails the drift check below recomputes them
        if (fp32_products(s)) launch_renorm_columns(s->phi, s->N, R, s->norms0, s->norms1, st);
        launch_outputs(s);
        DH_HIP(hipMemcpyAsync(n0.data(), s->norms0, sizeof(double) * R, hipMemcpyDeviceToHost, st));
        DH_HIP(hipMemcpyAsync(n1.data(), s->norms1, sizeof(double) * R, hipMemcpyDeviceToHost, st));
        DH_HIP(hipStreamSynchronize(st));
        double drift = 0.0;
        for (int c = 0; c < R; ++c) {
            if (n0[c] > 0) drift = std::max(drift, std::fabs(n1[c] - n0[c]) / n0[c]);
            if (!std::isfinite(n1[c])) drift = INFINITY;
        }
        s->stats[4] = drift;
        if (s->timing) {
            float ms = 0.0f;
            DH_HIP(hipEventElapsedTime(&ms, s->ev0, s->ev1));
            s->stats[5] += ms;
            s->stats[6] += s->K;
        }
        if (drift < drift_tol) {
            s->stats[0] = s->K + 1;
            s->stats[1] += 1;
            s->stats[2] += s->K + 3;
            if (n_terms) *n_terms = s->K + 1;
            return DH_OK;
        }
        // the bounds did not enclose the spectrum: restore and widen
        std::swap(s->phi, s->acc);
        s->widen *= 1.25;
        s->stats[3] += 1;
    }
    // outputs of the restored (pre-step) state, as before the call
    launch_outputs(s);
    DH_HIP(hipStreamSynchronize(st));
    dh::set_error("dh_propagate: Chebyshev series lost the norm (drift %g) after widening the spectral bounds",
                  s->stats[4]);
    return DH_ERR_DIVERGED;
}

static int check_ready(dh_system* s, const char* who) {
    if (!s->ops_ready) {
        dh::set_error("%s: operators not built", who);
        return DH_ERR_STATE;
    }
    if (!s->state_set) {
        dh::set_error("%s: wavepackets not set", who);
        return DH_ERR_STATE;
    }
    return DH_OK;
}

extern "C" int dh_propagate(dh_system* s, double t_init, double t_max, int* n_terms) {
    CHECK_ARG(s, "dh_propagate: null system");
    int rc = check_ready(s, "dh_propagate");
    if (rc) return rc;
    DH_HIP(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    const double dt = (t_max - t_init) / H_BAR;
    if (s->phi_dirty) {
        convert_ao_to_phi(s);
        s->phi_dirty = false;
    }
    if (!(dt > 0.0)) {
        launch_outputs(s);
        DH_HIP(hipStreamSynchronize(st));
        if (n_terms) *n_terms = 0;
        return DH_OK;
    }
    return propagate_checked(s, dt, n_terms);
}

// Operators from the coordinates already in s->coords (device): S and h, then
// the orthogonalisation (dh_build_operators after dh_set_coords).
// Complex elements of partial sums a matvec of R columns at N may need, over
// every kernel variant the autotuner can pick for the shape.
static long matvec_partial_cplx(int N, int R) {
    const int ld = round_ld(N);
    long part = (long)matvec_split_for(N, R) * ld * std::min(R, 16);  // split-K column-block variants
    if (sym_supported(R, TRI_FULL)) part = std::max(part, std::max(sym_partial_cplx(ld, R), sym2_partial_cplx(ld, R)));
    return std::max(part, (gemm3_partial_doubles(N, ld, R) + 1) / 2);
}

// Side stream, second workspace and the system's shadow operator set (allocated
// once, kept for later calls).
static int ensure_side(dh_system* s) {
    dh_context* c = s->ctx;
    if (!c->side) DH_HIP(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    if (c->ws2_ld < c->ws_ld) {
        free_ws2(c);
        const size_t n = (size_t)c->ws_ld * c->ws_ld;
        DALLOC(c->S2, n);
        DALLOC(c->h2, n);
        DALLOC(c->W2, n);
        DALLOC(c->Dinv2, (size_t)c->ws_ld * TILE);
        DALLOC(c->info2, 1);
        c->ws2_ld = c->ws_ld;
    }
    auto& h = s->sh;
    const size_t mat = (size_t)s->ld * s->ld;
    if (!h.allocated) {
        DALLOC(h.Q1, mat);
        DALLOC(h.Q2, mat);
        DALLOC(h.Hth, mat);
        DALLOC(h.coords, (size_t)s->n_atoms * 3);
        DALLOC(h.lv, (size_t)s->ld);
        DALLOC(h.lvp, (size_t)s->ld);
        DALLOC(h.lw, (size_t)s->ld);
        DALLOC(h.lal, LANCZOS_MAX + 2);
        DALLOC(h.lbe, LANCZOS_MAX + 2);
        DALLOC(h.partial, (size_t)matvec_partial_cplx(s->N, 1));
        DALLOC(h.counters, (size_t)s->ld);
        launch_zero_u32(h.counters, s->ld, c->side);
        launch_zero_cplx(h.lw, s->ld, c->side);
        DH_HIP(hipHostMalloc((void**)&h.h_info, sizeof(int), hipHostMallocDefault));
        DH_HIP(hipHostMalloc((void**)&h.h_al, sizeof(double) * (LANCZOS_MAX + 2), hipHostMallocDefault));
        DH_HIP(hipHostMalloc((void**)&h.h_be, sizeof(double) * (LANCZOS_MAX + 2), hipHostMallocDefault));
        DH_HIP(hipEventCreateWithFlags(&h.done, hipEventDisableTiming));
        h.allocated = true;
    }
    if (s->precision != 0 && !h.H32) DALLOC(h.H32, mat);
    if (s->precision == 3 && !h.Hlo) DALLOC(h.Hlo, mat);
    return DH_OK;
}

// The operators of the geometry at d_geo (device, n_atoms x 3) into the shadow
// set, on the side stream: S and h, Cholesky, L^-1, H_orth, the precision
// mode's FP32 copies and the Lanczos coefficients.  Only launches: the host
// does not wait here, so it goes on to queue the main stream's steps, which
// run meanwhile; side_finish completes the build.
static int side_launch(dh_system* s, const double* d_geo, double cutoff, double s_drop) {
    const auto t0 = std::chrono::steady_clock::now();
    dh_context* c = s->ctx;
    auto& h = s->sh;
    DH_HIP(hipMemcpyAsync(h.coords, d_geo, sizeof(double) * 3 * s->n_atoms, hipMemcpyDeviceToDevice, c->side));
    BasisDev b2 = s->bd;
    b2.coords = h.coords;
    if (launch_build_SH(b2, s->N, s->ld, c->S2, c->h2, cutoff, s_drop, c->side)) {
        dh::set_error("dh_propagate_steps: kernel launch failed: %s", hipGetErrorString(hipGetLastError()));
        return DH_ERR_HIP;
    }
    OpsTarget t{c->side, c->S2, c->h2, c->W2, c->Dinv2, c->info2, h.Q1, h.Q2, h.Hth, h.H32, h.Hlo,
                h.lv, h.lvp, h.lw, h.partial, h.lal, h.lbe, h.counters, h.bounds};
    int rc = factor_into(s, t, false, h.h_info);
    if (rc) return rc;
    f32_copies_into(s, t);
    rc = bounds_into(s, t, h.h_al, h.h_be, &h.m);
    if (rc) return rc;
    DH_HIP(hipEventRecord(h.done, c->side));
    h.pending = true;
    s->stats[8] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return DH_OK;
}

// Waits for the launched build (the side stream is idle afterwards), checks its
// Cholesky status and turns its Lanczos coefficients into the shadow bounds.
static int side_finish(dh_system* s) {
    auto& h = s->sh;
    if (!h.pending) return DH_OK;
    const auto t0 = std::chrono::steady_clock::now();
    h.pending = false;
    DH_HIP(hipEventSynchronize(h.done));
    s->stats[8] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (*h.h_info != 0) {
        dh::set_error("dh_propagate_steps: S of a rebuilt geometry is not positive definite (Cholesky failed at "
                      "column %d)", *h.h_info);
        return DH_ERR_NOT_SPD;
    }
    lanczos_bounds_host(h.h_al, h.h_be, h.m, s->rel_margin, h.bounds);
    return DH_OK;
}

// The shadow operators become the system's (host pointer swaps; both streams
// are idle): as after dh_build_operators on the new geometry.
static void side_commit(dh_system* s) {
    dh_context* c = s->ctx;
    auto& h = s->sh;
    std::swap(s->Q1, h.Q1);
    std::swap(s->Q2, h.Q2);
    std::swap(s->Hth, h.Hth);
    std::swap(s->H32, h.H32);
    std::swap(s->Hlo, h.Hlo);
    std::swap(s->coords, h.coords);
    s->bd.coords = s->coords;
    std::swap(c->S, c->S2);
    std::swap(c->h, c->h2);
    std::swap(c->W, c->W2);
    std::swap(c->Dinv, c->Dinv2);
    c->ws_owner = s;
    take_bounds(s, h.bounds);
    s->have_coords = true;
    s->ops_ready = true;
    s->phi_dirty = true;
    s->w_ready = true;
    s->hprime_ready = false;
    s->h32_ready = s->precision != 0;
    s->stats[7] += 1;
}

// Steps of one multi-step series starting now, at most `left`: the system's
// steps_per_series (0: MULTI_MAX), within the coefficient table (the Bessel
// recurrence of r m dt terms must start below KMAX_ALLOC).  1 means a series
// per step.
static int multi_chunk_steps(const dh_system* s, double dt, int left) {
    const int cap = s->steps_per_series > 0 ? std::min(s->steps_per_series, MULTI_MAX) : MULTI_MAX;
    int m = std::min(left, cap);
    const double r = 0.5 * (s->emax - s->emin) * s->widen;
    while (m > 1) {
        const double x = r * m * dt;
        if (x + 15.0 * std::cbrt(x) + 40.0 < KMAX_ALLOC - 2) break;
        --m;
    }
    return std::max(m, 1);
}

// The coefficient rows of an m-step series of s (host recurrence, one copy on
// the stream): accumulator q's row in s->coefm, its last term Kq[q] (made
// non-decreasing in q, a shorter row padded with zero coefficients: adding
// 0 T_k changes nothing) in Kq and s->kq_d.  The buffers are allocated on
// first use.
static int multi_coefficients(dh_system* s, double dt, int m, int* Kq) {
    hipStream_t st = s->ctx->stream;
    const long acc_n = (long)s->ld * 2 * s->W;
    constexpr int CS = KMAX_ALLOC + 2;
    if (!s->accm) {
        DALLOC(s->accm, (size_t)MULTI_MAX * acc_n);
        // the ring's rows N..ld-1 are never written and must read as finite
        // inputs of the next term
        DALLOC(s->ring, (size_t)PANEL * acc_n);
        DH_HIP(hipMemsetAsync(s->ring, 0, sizeof(cplx) * (size_t)PANEL * acc_n, st));
        DALLOC(s->coefm, (size_t)MULTI_MAX * CS);
        DALLOC(s->kq_d, (size_t)MULTI_MAX);
        DH_HIP(hipHostMalloc((void**)&s->coefm_h, sizeof(cplx) * MULTI_MAX * CS, hipHostMallocDefault));
        DH_HIP(hipHostMalloc((void**)&s->kq_h, sizeof(int) * MULTI_MAX, hipHostMallocDefault));
        DH_HIP(hipEventCreateWithFlags(&s->coefm_ev, hipEventDisableTiming));
    }
    if (s->coefm_pending) DH_HIP(hipEventSynchronize(s->coefm_ev));  // the staging buffers are free again
    const double c = 0.5 * (s->emax + s->emin);
    const double r = 0.5 * (s->emax - s->emin) * s->widen;
    for (int q = 0; q < m; ++q) {
        cplx* row = s->coefm_h + (size_t)q * CS;
        const int k = bessel_coef_host(r * (q + 1) * dt, c * (q + 1) * dt, KMAX_ALLOC, BESSEL_TOL, row);
        if (k <= 0 && r * dt > 1e-12) {
            dh::set_error("Chebyshev coefficients: no terms for r*dt = %g", r * (q + 1) * dt);
            return DH_ERR_STATE;
        }
        Kq[q] = k;
    }
    for (int q = 1; q < m; ++q)
        if (Kq[q] < Kq[q - 1]) {
            cplx* row = s->coefm_h + (size_t)q * CS;
            for (int k = Kq[q] + 1; k <= Kq[q - 1] + 1; ++k) row[k] = cplx{0.0, 0.0};
            Kq[q] = Kq[q - 1];
        }
    for (int q = 0; q < m; ++q) s->kq_h[q] = Kq[q];
    DH_HIP(hipMemcpyAsync(s->coefm, s->coefm_h, sizeof(cplx) * (size_t)m * CS, hipMemcpyHostToDevice, st));
    DH_HIP(hipMemcpyAsync(s->kq_d, s->kq_h, sizeof(int) * m, hipMemcpyHostToDevice, st));
    DH_HIP(hipEventRecord(s->coefm_ev, st));
    s->coefm_pending = true;
    return DH_OK;
}

// m steps of dt from s->phi in one Chebyshev recurrence: the terms T_k(Hh) phi
// do not depend on time, so step q's state sum_k a_k((q + 1) dt) T_k phi
// collects in accumulator q of s->accm with its own coefficients, all from the
// same T_k.  The series runs to the last step's K; accumulator q stops taking
// terms after its own K_q.  Results equal m single steps to rounding (the
// propagator is exact in both), with about r^(1/3) (m - m^(1/3)) fewer terms
// than m series of one step.  Returns the terms of the series in *kmax.
static int multi_series(dh_system* s, double dt, int m, int* kmax) {
    hipStream_t st = s->ctx->stream;
    const int R = 2 * s->W;
    const long acc_n = (long)s->ld * R;
    constexpr int CS = KMAX_ALLOC + 2;
    int Kq[MULTI_MAX];
    if (int rc = multi_coefficients(s, dt, m, Kq)) return rc;
    const double c = 0.5 * (s->emax + s->emin);
    const double r = 0.5 * (s->emax - s->emin) * s->widen;
    SeriesOps o{};
    o.H = s->Hth;
    o.H32 = s->H32;
    o.Hlo = s->Hlo;
    o.ld = s->ld;
    o.N = s->N;
    o.R = R;
    o.phi = s->phi;
    o.x0 = s->x0;
    o.x1 = s->x1;
    o.partial = s->partial;
    o.counters = s->counters;
    o.x32[0] = s->x32[0];
    o.x32[1] = s->x32[1];
    o.xb16[0] = s->xb16[0];
    o.xb16[1] = s->xb16[1];
    o.lib = s->ctx->lib;
    o.prec = s->precision;
    o.c = c;
    o.r = r;
    o.m = m;
    o.accm = s->accm;
    o.accm_stride = acc_n;
    o.coefm = s->coefm;
    o.coefm_stride = CS;
    o.kq = s->kq_d;
    o.Kq = Kq;
    o.ring = s->ring;
    o.ring_stride = acc_n;
    o.K = Kq[m - 1];
    if (int rc = series_core(o, st)) return rc;
    DH_HIP(hipGetLastError());
    *kmax = Kq[m - 1];
    return DH_OK;
}

// Device buffers of one dh_propagate_steps call, freed on every return path
// (after the side stream, whose launched build reads geo, has drained).
struct StepsScratch {
    double *geo = nullptr, *n0 = nullptr, *n1 = nullptr, *pops = nullptr, *erg = nullptr;
    cplx *ck_phi = nullptr, *ck_out = nullptr;
    hipStream_t side = nullptr;
    ~StepsScratch() {
        if (side) hipStreamSynchronize(side);
        for (void* p : {(void*)geo, (void*)n0, (void*)n1, (void*)pops, (void*)erg, (void*)ck_phi, (void*)ck_out})
            if (p) hipFree(p);
    }
};

extern "C" int dh_propagate_steps(dh_system* s, int n_steps, double dt_ps, int rebuild_every, const double* coords,
                                  double cutoff, double s_drop, double* pops, double* erg, int* n_terms) {
    CHECK_ARG(s, "dh_propagate_steps: null system");
    CHECK_ARG(n_steps > 0 && dt_ps > 0.0 && rebuild_every >= 0, "dh_propagate_steps: need n_steps > 0, dt > 0, "
              "rebuild_every >= 0");
    const int n_geo = rebuild_every > 0 ? (n_steps - 1) / rebuild_every : 0;
    CHECK_ARG(n_geo == 0 || (coords && s->n_atoms > 0), "dh_propagate_steps: rebuilds need coordinates of a system "
              "built from atoms");
    int rc = check_ready(s, "dh_propagate_steps");
    if (rc) return rc;
    DH_HIP(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    const int W = s->W, R = 2 * W, NP = s->n_frag + 1;
    const double dt = dt_ps / H_BAR;
    StepsScratch d;
    DALLOC(d.n0, (size_t)n_steps * R);
    DALLOC(d.n1, (size_t)n_steps * R);
    DALLOC(d.pops, (size_t)n_steps * W * NP);
    DALLOC(d.erg, (size_t)n_steps * W * 2);
    DALLOC(d.ck_phi, (size_t)s->ld * R);
    DALLOC(d.ck_out, (size_t)s->ld * 3 * W);
    const size_t geo_n = (size_t)s->n_atoms * 3;
    if (n_geo > 0) {
        if ((rc = ensure_side(s))) return rc;
        d.side = s->ctx->side;
        DALLOC(d.geo, geo_n * n_geo);
        // every geometry goes to the device once; a rebuild reads its slice there
        DH_HIP(hipMemcpyAsync(d.geo, coords, sizeof(double) * geo_n * n_geo, hipMemcpyHostToDevice, st));
        DH_HIP(hipStreamSynchronize(st));  // the side stream reads them
    }
    std::vector<double> h0((size_t)n_steps * R), h1((size_t)n_steps * R);
    const double drift_tol = fp32_products(s) ? 1e-4 : 1e-9;
    int last_terms = 0;
    for (int j = 0; j < n_steps;) {
        // the operators of this step's geometry were built while the previous
        // interval's steps ran
        if (j > 0 && rebuild_every > 0 && j % rebuild_every == 0) {
            if ((rc = side_finish(s))) return rc;
            side_commit(s);
        }
        const int jend = rebuild_every > 0 ? std::min(n_steps, (j / rebuild_every + 1) * rebuild_every) : n_steps;
        const bool build_next = rebuild_every > 0 && jend < n_steps && jend % rebuild_every == 0;
        if (s->phi_dirty) {
            convert_ao_to_phi(s);
            s->phi_dirty = false;
        }
        // state at the start of the interval, for a redo of its steps one by one
        DH_HIP(hipMemcpyAsync(d.ck_phi, s->phi, sizeof(cplx) * s->ld * R, hipMemcpyDeviceToDevice, st));
        DH_HIP(hipMemcpyAsync(d.ck_out, s->out, sizeof(cplx) * s->ld * 3 * W, hipMemcpyDeviceToDevice, st));
        if ((rc = ensure_coefficients(s, dt))) return rc;
        if (s->timing)
            while (s->evpool.size() < 2 * (size_t)(jend - j)) {
                hipEvent_t e = nullptr;
                DH_HIP(hipEventCreate(&e));
                s->evpool.push_back(e);
            }
        // the steps run as multi-step series (up to MULTI_MAX steps per
        // recurrence) in every precision mode and with every term kernel: the
        // accumulators take the terms in panel passes, outside the term kernels
        const bool multi_ok = s->steps_per_series != 1;
        int n_timed = 0, chunk_terms = 0;
        long terms_timed = 0, launches = 0;
        for (int i = j; i < jend;) {
            const int mc = multi_ok ? multi_chunk_steps(s, dt, jend - i) : 1;
            if (mc >= 2) {
                if (s->timing) DH_HIP(hipEventRecord(s->evpool[2 * n_timed], st));
                int kmax = 0;
                if ((rc = multi_series(s, dt, mc, &kmax))) return rc;
                if (s->timing) DH_HIP(hipEventRecord(s->evpool[2 * n_timed + 1], st));
                if (s->timing) {
                    ++n_timed;
                    terms_timed += kmax;
                }
                launches += kmax + 3L * mc;
                chunk_terms = kmax + 1;
                const long acc_n = (long)s->ld * R;
                cplx* keep = s->phi;
                for (int q = 0; q < mc; ++q) {
                    cplx* st_q = s->accm + (size_t)q * acc_n;
                    // FP32 products: every step's state takes back the norm of the
                    // series' input (as the single steps restore each step's), so
                    // the drift checked is that of q + 1 steps
                    const bool f32p = fp32_products(s);
                    launch_column_norms(s->N, R, (q == 0 || f32p) ? keep : st_q - acc_n, d.n0 + (size_t)(i + q) * R, st);
                    launch_column_norms(s->N, R, st_q, d.n1 + (size_t)(i + q) * R, st);
                    if (f32p) launch_renorm_columns(st_q, s->N, R, d.n0 + (size_t)(i + q) * R, d.n1 + (size_t)(i + q) * R, st);
                    s->phi = st_q;  // the outputs of step i + q
                    launch_outputs(s);
                    s->phi = keep;
                    DH_HIP(hipMemcpyAsync(d.pops + (size_t)(i + q) * W * NP, s->pops, sizeof(double) * W * NP,
                                          hipMemcpyDeviceToDevice, st));
                    DH_HIP(hipMemcpyAsync(d.erg + (size_t)(i + q) * W * 2, s->erg, sizeof(double) * W * 2,
                                          hipMemcpyDeviceToDevice, st));
                }
                DH_HIP(hipMemcpyAsync(s->phi, s->accm + (size_t)(mc - 1) * acc_n, sizeof(cplx) * acc_n,
                                      hipMemcpyDeviceToDevice, st));
            } else {
                launch_column_norms(s->N, R, s->phi, d.n0 + (size_t)i * R, st);
                if (s->timing) DH_HIP(hipEventRecord(s->evpool[2 * n_timed], st));
                if ((rc = chebyshev_series(s))) return rc;
                if (s->timing) DH_HIP(hipEventRecord(s->evpool[2 * n_timed + 1], st));
                if (s->timing) {
                    ++n_timed;
                    terms_timed += s->K;
                }
                launches += s->K + 3L;
                chunk_terms = s->K + 1;
                std::swap(s->phi, s->acc);
                launch_column_norms(s->N, R, s->phi, d.n1 + (size_t)i * R, st);
                if (fp32_products(s))
                    launch_renorm_columns(s->phi, s->N, R, d.n0 + (size_t)i * R, d.n1 + (size_t)i * R, st);
                launch_outputs(s);
                DH_HIP(hipMemcpyAsync(d.pops + (size_t)i * W * NP, s->pops, sizeof(double) * W * NP,
                                      hipMemcpyDeviceToDevice, st));
                DH_HIP(hipMemcpyAsync(d.erg + (size_t)i * W * 2, s->erg, sizeof(double) * W * 2,
                                      hipMemcpyDeviceToDevice, st));
            }
            // the next geometry's operators, on the side stream, while this
            // interval's steps run; launched once the first series is queued so
            // that the main stream has work meanwhile (the shadow set and the
            // second workspace hold nothing the main stream still reads)
            if (i == j && build_next) {
                if ((rc = side_launch(s, d.geo + geo_n * (jend / rebuild_every - 1), cutoff, s_drop))) return rc;
            }
            i += mc;
        }
        DH_HIP(hipGetLastError());
        const size_t cnt = (size_t)(jend - j) * R;
        DH_HIP(hipMemcpyAsync(h0.data() + (size_t)j * R, d.n0 + (size_t)j * R, sizeof(double) * cnt,
                              hipMemcpyDeviceToHost, st));
        DH_HIP(hipMemcpyAsync(h1.data() + (size_t)j * R, d.n1 + (size_t)j * R, sizeof(double) * cnt,
                              hipMemcpyDeviceToHost, st));
        DH_HIP(hipStreamSynchronize(st));
        float series_ms = 0.0f;
        if (s->timing)
            for (int i = 0; i < n_timed; ++i) {
                float ms = 0.0f;
                DH_HIP(hipEventElapsedTime(&ms, s->evpool[2 * i], s->evpool[2 * i + 1]));
                series_ms += ms;
            }
        double drift = 0.0;
        for (size_t e = (size_t)j * R; e < (size_t)jend * R; ++e) {
            if (h0[e] > 0) drift = std::max(drift, std::fabs(h1[e] - h0[e]) / h0[e]);
            if (!std::isfinite(h1[e])) drift = INFINITY;
        }
        if (drift < drift_tol) {
            s->stats[4] = drift;
            s->stats[0] = chunk_terms;
            s->stats[1] += jend - j;
            s->stats[2] += (double)launches;
            if (s->timing) {
                s->stats[5] += series_ms;
                s->stats[6] += (double)terms_timed;
            }
            last_terms = chunk_terms;
        } else {
            // a step left the bounds: back to the interval's start, then step by
            // step with the checked step (wider bounds where needed)
            DH_HIP(hipMemcpyAsync(s->phi, d.ck_phi, sizeof(cplx) * s->ld * R, hipMemcpyDeviceToDevice, st));
            DH_HIP(hipMemcpyAsync(s->out, d.ck_out, sizeof(cplx) * s->ld * 3 * W, hipMemcpyDeviceToDevice, st));
            for (int i = j; i < jend; ++i) {
                if ((rc = propagate_checked(s, dt, &last_terms))) return rc;
                DH_HIP(hipMemcpyAsync(d.pops + (size_t)i * W * NP, s->pops, sizeof(double) * W * NP,
                                      hipMemcpyDeviceToDevice, st));
                DH_HIP(hipMemcpyAsync(d.erg + (size_t)i * W * 2, s->erg, sizeof(double) * W * 2,
                                      hipMemcpyDeviceToDevice, st));
            }
        }
        j = jend;
    }
    if (pops) DH_HIP(hipMemcpyAsync(pops, d.pops, sizeof(double) * n_steps * W * NP, hipMemcpyDeviceToHost, st));
    if (erg) DH_HIP(hipMemcpyAsync(erg, d.erg, sizeof(double) * n_steps * W * 2, hipMemcpyDeviceToHost, st));
    DH_HIP(hipStreamSynchronize(st));
    if (n_terms) *n_terms = last_terms;
    return DH_OK;
}

extern "C" int dh_get_observables(dh_system* s, int wp, double* ao_bra, double* pops, double* erg) {
    CHECK_ARG(s && wp >= 0 && wp < s->W, "dh_get_observables: bad argument");
    DH_HIP(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    const int W = s->W;
    if (ao_bra) {
        const int rc = get_column(s, ao_bra, s->out + W + wp, 3 * W, 0);
        if (rc) return rc;
    }
    if (pops)
        DH_HIP(hipMemcpyAsync(pops, s->pops + (size_t)wp * (s->n_frag + 1), sizeof(double) * (s->n_frag + 1),
                              hipMemcpyDeviceToHost, st));
    if (erg) DH_HIP(hipMemcpyAsync(erg, s->erg + 2 * wp, sizeof(double) * 2, hipMemcpyDeviceToHost, st));
    DH_HIP(hipStreamSynchronize(st));
    return DH_OK;
}

extern "C" int dh_set_graph(dh_system* s, int on) {
    CHECK_ARG(s, "dh_set_graph: null system");
    s->use_graph = on != 0;
    return DH_OK;
}

extern "C" int dh_get_stats(dh_system* s, double* out9) {
    CHECK_ARG(s && out9, "dh_get_stats: null argument");
    for (int i = 0; i < 9; ++i) out9[i] = s->stats[i];
    return DH_OK;
}

extern "C" int dh_set_timing(dh_system* s, int on) {
    CHECK_ARG(s, "dh_set_timing: null system");
    s->timing = on != 0;
    return DH_OK;
}

extern "C" int dh_set_steps_per_series(dh_system* s, int m) {
    CHECK_ARG(s && m >= 0, "dh_set_steps_per_series: need a system and m >= 0");
    s->steps_per_series = m;
    return DH_OK;
}

extern "C" int dh_propagation_elhl(dh_context* ctx, const int* Np, const double* S, const double* h, double* H_out,
                                   double* AO_bra, double* AO_ket, double* PSI_bra, double* PSI_ket,
                                   const double* t_init, const double* t_max, double* tau, double* save_tau) {
    (void)AO_ket;
    CHECK_ARG(ctx && Np && S && h && PSI_bra && PSI_ket && t_init && t_max, "dh_propagation_elhl: null argument");
    const int N = *Np;
    CHECK_ARG(N > 0, "dh_propagation_elhl: N <= 0");
    if (ctx->dropin && ctx->dropin_N != N) {
        dh_system_destroy(ctx->dropin);
        ctx->dropin = nullptr;
    }
    int rc;
    if (!ctx->dropin) {
        rc = dh_system_create(ctx, 0, nullptr, nullptr, N, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                              nullptr, nullptr, nullptr, nullptr, 1, 1, &ctx->dropin);
        if (rc) return rc;
        ctx->dropin_N = N;
    }
    dh_system* s = ctx->dropin;
    if ((rc = dh_load_operators(s, S, h, N))) return rc;
    if ((rc = dh_orthogonalize(s, H_out != nullptr))) return rc;
    if (H_out && (rc = dh_get_matrix(s, DH_MAT_HPRIME, H_out, N))) return rc;
    if ((rc = dh_set_wavepacket(s, 0, PSI_bra, PSI_ket))) return rc;
    if ((rc = dh_propagate(s, *t_init, *t_max, nullptr))) return rc;
    if ((rc = dh_get_wavepacket(s, 0, PSI_bra, PSI_ket))) return rc;
    if (AO_bra && (rc = dh_get_observables(s, 0, AO_bra, nullptr, nullptr))) return rc;
    const double step = (*t_max - *t_init) / H_BAR;
    if (save_tau) *save_tau = step;
    if (tau) *tau = step;
    return DH_OK;
}

// ------------------------------------------------------------------ caller-owned device operands
// Work layout of dh_chebyshev_apply: recurrence vectors, the term kernels'
// partial sums (enough for every variant the autotuner may pick for (N, R)),
// counters, Bessel table and coefficients, norms; 256-byte aligned pieces.
struct ApplyWork {
    cplx *x0, *x1, *acc, *partial;
    unsigned* counters;
    double *J, *norms0, *norms1;
    cplx* coef;
    int* kout;
};

static size_t apply_layout(int N, int R, char* base, ApplyWork* w) {
    const int ld = round_ld(N);
    long part = (long)matvec_split_for(N, R) * ld * R;  // split-K column-block variants
    if (sym_supported(R, TRI_FULL)) part = std::max(part, std::max(sym_partial_cplx(ld, R), sym2_partial_cplx(ld, R)));
    part = std::max(part, (gemm3_partial_doubles(N, ld, R) + 1) / 2);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char* p = base ? base + off : nullptr;
        off += (bytes + 255) & ~size_t(255);
        return p;
    };
    ApplyWork t{};
    t.x0 = (cplx*)take(sizeof(cplx) * ld * R);
    t.x1 = (cplx*)take(sizeof(cplx) * ld * R);
    t.acc = (cplx*)take(sizeof(cplx) * ld * R);
    t.partial = (cplx*)take(sizeof(cplx) * part);
    t.counters = (unsigned*)take(sizeof(unsigned) * ld);
    t.J = (double*)take(sizeof(double) * (KMAX_ALLOC + 2));
    t.coef = (cplx*)take(sizeof(cplx) * (KMAX_ALLOC + 2));
    t.kout = (int*)take(sizeof(int) * 4);
    t.norms0 = (double*)take(sizeof(double) * R);
    t.norms1 = (double*)take(sizeof(double) * R);
    if (w) *w = t;
    return off;
}

extern "C" long long dh_chebyshev_work_bytes(int N, int R) {
    if (N <= 0 || !(R == 1 || R == 2 || R == 4 || R == 8)) return -1;
    return (long long)apply_layout(N, R, nullptr, nullptr);
}

// phi <- exp(-i H dt / hbar) phi on the context stream, with the work buffer
// carved as apply_layout gives it.  Shared by the workspace and the
// allocation-free entry points, so both run the same kernels on the same shape.
static int chebyshev_apply_impl(dh_context* ctx, int N, const double* H, int ldh, const double* bounds, double dt_ps,
                                int R, double* phi_d, char* work, int* n_terms) {
    hipStream_t st = ctx->stream;
    ApplyWork w{};
    apply_layout(N, R, work, &w);
    cplx* phi = reinterpret_cast<cplx*>(phi_d);
    // rows N..ld-1 of the state are padding: zero, so they stay zero under H
    if (ldh > N) launch_zero_cplx(phi + (size_t)N * R, (long)(ldh - N) * R, st);
    launch_zero_u32(w.counters, ldh, st);
    const double dt = dt_ps / H_BAR;
    if (!(dt > 0.0)) {
        if (n_terms) *n_terms = 0;
        DH_HIP(hipGetLastError());
        DH_HIP(hipStreamSynchronize(st));
        return DH_OK;
    }
    launch_column_norms(N, R, phi, w.norms0, st);
    SeriesOps o{};
    o.H = H;
    o.ld = ldh;
    o.N = N;
    o.R = R;
    o.phi = phi;
    o.x0 = w.x0;
    o.x1 = w.x1;
    o.acc = w.acc;
    o.partial = w.partial;
    o.counters = w.counters;
    o.coef = w.coef;
    o.prec = 0;
    double widen = 1.0, drift = 0.0;
    double n0[8], n1[8];  // R <= 8 (check_apply_args); no host allocation on this path
    for (int attempt = 0; attempt < 4; ++attempt) {
        const double c = 0.5 * (bounds[1] + bounds[0]);
        const double r = 0.5 * (bounds[1] - bounds[0]) * widen;
        launch_bessel(r * dt, c * dt, KMAX_ALLOC, BESSEL_TOL, w.J, w.coef, w.kout, st);
        int K = 0;
        DH_HIP(hipGetLastError());
        DH_HIP(hipMemcpyAsync(&K, w.kout, sizeof(int), hipMemcpyDeviceToHost, st));
        DH_HIP(hipStreamSynchronize(st));
        if (K <= 0 && r * dt > 1e-12) {
            dh::set_error("dh_chebyshev_apply: no Chebyshev terms for r*dt = %g", r * dt);
            return DH_ERR_STATE;
        }
        o.K = K;
        o.c = c;
        o.r = r;
        if (int rc = series_core(o, st)) return rc;
        launch_column_norms(N, R, w.acc, w.norms1, st);
        DH_HIP(hipGetLastError());
        DH_HIP(hipMemcpyAsync(n0, w.norms0, sizeof(double) * R, hipMemcpyDeviceToHost, st));
        DH_HIP(hipMemcpyAsync(n1, w.norms1, sizeof(double) * R, hipMemcpyDeviceToHost, st));
        DH_HIP(hipStreamSynchronize(st));
        drift = 0.0;
        for (int col = 0; col < R; ++col) {
            if (n0[col] > 0) drift = std::max(drift, std::fabs(n1[col] - n0[col]) / n0[col]);
            if (!std::isfinite(n1[col])) drift = INFINITY;
        }
        if (drift < 1e-9) {
            DH_HIP(hipMemcpyAsync(phi, w.acc, sizeof(cplx) * (size_t)N * R, hipMemcpyDeviceToDevice, st));
            DH_HIP(hipStreamSynchronize(st));
            if (n_terms) *n_terms = K + 1;
            return DH_OK;
        }
        widen *= 1.25;  // the bounds did not enclose the spectrum; phi is untouched
    }
    dh::set_error("dh_chebyshev_apply: Chebyshev series lost the norm (drift %g) after widening the bounds", drift);
    return DH_ERR_DIVERGED;
}

static int check_apply_args(const char* who, dh_context* ctx, int N, const double* H, int ldh, const double* bounds,
                            int R, const double* phi) {
    if (!(ctx && H && bounds && phi)) {
        dh::set_error("%s: null argument", who);
        return DH_ERR_ARG;
    }
    if (N <= 0 || ldh < N || ldh % TILE != 0 || ldh != round_ld(N)) {
        dh::set_error("%s: need N > 0 and ldh = N rounded up to a multiple of %d (got N = %d, ldh = %d)", who, TILE, N,
                      ldh);
        return DH_ERR_ARG;
    }
    if (!(R == 1 || R == 2 || R == 4 || R == 8)) {
        dh::set_error("%s: R must be 1, 2, 4 or 8 (got %d)", who, R);
        return DH_ERR_ARG;
    }
    if (!(bounds[1] > bounds[0]) || !std::isfinite(bounds[0]) || !std::isfinite(bounds[1])) {
        dh::set_error("%s: bounds must be finite with E_max > E_min", who);
        return DH_ERR_ARG;
    }
    return DH_OK;
}

extern "C" int dh_chebyshev_apply(dh_context* ctx, int N, const double* H, int ldh, const double* bounds, double dt,
                                  int R, double* phi, void* work, long long work_bytes, int* n_terms) {
    int rc = check_apply_args("dh_chebyshev_apply", ctx, N, H, ldh, bounds, R, phi);
    if (rc) return rc;
    const long long need = dh_chebyshev_work_bytes(N, R);
    if (!work || work_bytes < need || (reinterpret_cast<uintptr_t>(work) & 255)) {
        dh::set_error("dh_chebyshev_apply: work must be a 256-byte aligned device buffer of at least %lld bytes "
                      "(dh_chebyshev_work_bytes)", need);
        return DH_ERR_ARG;
    }
    DH_HIP(hipSetDevice(ctx->device));
    return chebyshev_apply_impl(ctx, N, H, ldh, bounds, dt, R, phi, static_cast<char*>(work), n_terms);
}

extern "C" int dh_chebyshev_apply_noalloc(dh_context* ctx, int N, const double* H, int ldh, const double* bounds,
                                          double dt, int R, double* phi, int* n_terms) {
    int rc = check_apply_args("dh_chebyshev_apply_noalloc", ctx, N, H, ldh, bounds, R, phi);
    if (rc) return rc;
    const long long need = dh_chebyshev_work_bytes(N, R);
    if (need > NOALLOC_SCRATCH_BYTES) {
        dh::set_error("dh_chebyshev_apply_noalloc: N = %d, R = %d needs %lld bytes of scratch, beyond the %lld the "
                      "library holds (N <= 4096 with R <= 4, or N <= 2048 with R <= 8); use dh_chebyshev_apply",
                      N, R, need, NOALLOC_SCRATCH_BYTES);
        return DH_ERR_ARG;
    }
    void* ptr = (ctx->device >= 0 && ctx->device < MAX_DEVICES) ? g_scratch_dev[ctx->device] : nullptr;
    if (!ptr) {
        dh::set_error("dh_chebyshev_apply_noalloc: no scratch address for device %d", ctx->device);
        return DH_ERR_STATE;
    }
    DH_HIP(hipSetDevice(ctx->device));
    return chebyshev_apply_impl(ctx, N, H, ldh, bounds, dt, R, phi, static_cast<char*>(ptr), n_terms);
}

// ------------------------------------------------------------------ test hooks
#include "../../include/dynemol_hip_testing.h"

extern "C" int dh_test_device_alloc(long long bytes, void** ptr) {
    CHECK_ARG(bytes > 0 && ptr, "dh_test_device_alloc: bad argument");
    char* p = nullptr;
    DALLOC(p, (size_t)bytes);
    *ptr = p;
    return DH_OK;
}

extern "C" int dh_test_device_free(void* ptr) {
    DH_HIP(hipFree(ptr));
    return DH_OK;
}

extern "C" int dh_test_memcpy(void* dst, const void* src, long long bytes, int to_device) {
    CHECK_ARG(dst && src && bytes >= 0, "dh_test_memcpy: bad argument");
    DH_HIP(hipMemcpy(dst, src, (size_t)bytes, to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost));
    return DH_OK;
}

extern "C" int dh_test_gemm(int M, int N, int K, const double* A, int lda, const double* B, int ldb, double* C,
                            int ldc, double alpha, double beta, int transA, int transB, int kmode, int lower_only) {
    CHECK_ARG(M > 0 && N > 0 && K > 0 && A && B && C, "dh_test_gemm: bad argument");
    const long na = (long)lda * (transA ? M : K), nb = (long)ldb * (transB ? K : N), nc = (long)ldc * N;
    double *dA = nullptr, *dB = nullptr, *dC = nullptr;
    DALLOC(dA, na);
    DALLOC(dB, nb);
    DALLOC(dC, nc);
    DH_HIP(hipMemcpy(dA, A, sizeof(double) * na, hipMemcpyHostToDevice));
    DH_HIP(hipMemcpy(dB, B, sizeof(double) * nb, hipMemcpyHostToDevice));
    DH_HIP(hipMemcpy(dC, C, sizeof(double) * nc, hipMemcpyHostToDevice));
    Gemm g{};
    g.M = M; g.N = N; g.K = K;
    g.A = dA; g.lda = lda; g.B = dB; g.ldb = ldb; g.C = dC; g.ldc = ldc;
    g.alpha = alpha; g.beta = beta; g.transA = transA; g.transB = transB; g.kmode = kmode; g.lower_only = lower_only;
    g.batch = 1;
    launch_gemm(g, nullptr);
    DH_HIP(hipGetLastError());
    DH_HIP(hipDeviceSynchronize());
    DH_HIP(hipMemcpy(C, dC, sizeof(double) * nc, hipMemcpyDeviceToHost));
    hipFree(dA); hipFree(dB); hipFree(dC);
    return DH_OK;
}

extern "C" int dh_test_cholesky(int n, double* A, double* Linv, int* info) {
    CHECK_ARG(n > 0 && n % TILE == 0 && A && Linv && info, "dh_test_cholesky: bad argument");
    const long nn = (long)n * n;
    double *dA = nullptr, *dX = nullptr, *dD = nullptr, *dw = nullptr;
    int* di = nullptr;
    DALLOC(dA, nn);
    DALLOC(dX, nn);
    DALLOC(dD, (long)n * TILE);
    DALLOC(dw, nn / 2 + TILE);
    DALLOC(di, 1);
    DH_HIP(hipMemcpy(dA, A, sizeof(double) * nn, hipMemcpyHostToDevice));
    DH_HIP(hipMemset(di, 0, sizeof(int)));
    launch_cholesky(dA, n, dD, di, nullptr);
    launch_trtri_lower(dA, dX, dD, n, dw, nullptr);
    DH_HIP(hipGetLastError());
    DH_HIP(hipDeviceSynchronize());
    DH_HIP(hipMemcpy(A, dA, sizeof(double) * nn, hipMemcpyDeviceToHost));
    DH_HIP(hipMemcpy(Linv, dX, sizeof(double) * nn, hipMemcpyDeviceToHost));
    DH_HIP(hipMemcpy(info, di, sizeof(int), hipMemcpyDeviceToHost));
    hipFree(dA); hipFree(dX); hipFree(dD); hipFree(dw); hipFree(di);
    return DH_OK;
}

__global__ void test_copy_kernel(const double* __restrict__ a, double* __restrict__ b, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = a[i];
}

extern "C" int dh_test_copy(long long bytes, int repeats, double* ms_per_copy) {
    CHECK_ARG(bytes >= 8 && repeats >= 1, "dh_test_copy: bad argument");
    const long n = (long)(bytes / 8);
    double *a = nullptr, *b = nullptr;
    DALLOC(a, n);
    DALLOC(b, n);
    DH_HIP(hipMemset(a, 0, sizeof(double) * n));
    hipEvent_t e0, e1;
    DH_HIP(hipEventCreate(&e0));
    DH_HIP(hipEventCreate(&e1));
    DH_HIP(hipEventRecord(e0, nullptr));
    for (int r = 0; r < repeats; ++r)
        hipLaunchKernelGGL(test_copy_kernel, dim3(4096), dim3(256), 0, nullptr, a, b, n);
    DH_HIP(hipEventRecord(e1, nullptr));
    DH_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    DH_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (ms_per_copy) *ms_per_copy = ms / repeats;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(a);
    hipFree(b);
    return DH_OK;
}

// Streaming read: each thread sums 16-byte loads over a grid-stride loop and
// the block writes one partial (the read ceiling a matvec over H can approach).
__global__ __launch_bounds__(256) void test_read_kernel(const double2* __restrict__ a, long n2, double* out) {
    double s = 0.0;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n2; i += (long)gridDim.x * 256L) {
        const double2 v = a[i];
        s += v.x + v.y;
    }
    __shared__ double sh[256];
    sh[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < 256; ++k) t += sh[k];
        out[blockIdx.x] = t;
    }
}

extern "C" int dh_test_bessel_coefficients(double x, double phase, double tol, int kmax, double* coef, int* K_out) {
    CHECK_ARG(x >= 0.0 && tol > 0.0 && kmax >= 4 && coef && K_out, "dh_test_bessel_coefficients: bad argument");
    *K_out = bessel_coef_host(x, phase, kmax, tol, reinterpret_cast<cplx*>(coef));
    return DH_OK;
}

extern "C" int dh_test_read(long long bytes, int repeats, double* ms_per_read) {
    CHECK_ARG(bytes >= 16 && repeats >= 1, "dh_test_read: bad argument");
    const long n2 = (long)(bytes / 16);
    const int grid = 2048;
    double2* a = nullptr;
    double* out = nullptr;
    DALLOC(a, n2);
    DALLOC(out, grid);
    DH_HIP(hipMemset(a, 0, sizeof(double2) * n2));
    hipEvent_t e0, e1;
    DH_HIP(hipEventCreate(&e0));
    DH_HIP(hipEventCreate(&e1));
    hipLaunchKernelGGL(test_read_kernel, dim3(grid), dim3(256), 0, nullptr, a, n2, out);
    DH_HIP(hipEventRecord(e0, nullptr));
    for (int r = 0; r < repeats; ++r)
        hipLaunchKernelGGL(test_read_kernel, dim3(grid), dim3(256), 0, nullptr, a, n2, out);
    DH_HIP(hipEventRecord(e1, nullptr));
    DH_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    DH_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (ms_per_read) *ms_per_read = ms / repeats;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(a);
    hipFree(out);
    return DH_OK;
}

extern "C" int dh_test_force_matvec_variant(int variant) {
    CHECK_ARG(variant >= -1 && variant < matvec_variant_count(), "dh_test_force_matvec_variant: bad variant");
    matvec_force_variant(variant);
    return DH_OK;
}

extern "C" int dh_matvec_tuning(dh_system* s, int* out7) {
    CHECK_ARG(s && out7, "dh_matvec_tuning: null argument");
    const int W = s->W, R = 2 * W;
    const int q[7][2] = {{R, TRI_FULL}, {W, TRI_FULL}, {W, TRI_UPPER}, {R, TRI_LOWER}, {W, TRI_LOWER}, {W, TRI_UPPER}, {1, TRI_FULL}};
    for (int k = 0; k < 7; ++k) out7[k] = matvec_variant_for(s->N, q[k][0], q[k][1]);
    return DH_OK;
}

extern "C" int dh_test_matvec_timing(int N, int R, int variant, int reps, double* us_per_launch) {
    CHECK_ARG(N > 0 && R >= 1 && R <= 128 && variant >= -2 && variant < matvec_variant_count() && reps > 0,
              "dh_test_matvec_timing: bad argument");
    const int ld = round_ld(N);
    double* M = nullptr;
    cplx *x = nullptr, *x2 = nullptr, *out = nullptr, *acc = nullptr, *coef = nullptr, *partial = nullptr;
    unsigned* counters = nullptr;
    const int split = matvec_split_for(N, R);
    DALLOC(M, (size_t)ld * ld);
    DALLOC(x, (size_t)ld * R);
    DALLOC(x2, (size_t)ld * R);
    DALLOC(out, (size_t)ld * R);
    DALLOC(acc, (size_t)ld * R);
    DALLOC(coef, 8);
    DALLOC(partial, (size_t)std::max(std::max((long)split * ld * 16, sym_partial_cplx(ld, R <= 8 ? R : 8)),
                                     std::max(sym2_partial_cplx(ld, R <= 8 ? R : 8),
                                              (gemm3_partial_doubles(N, ld, R) + 1) / 2)));
    DALLOC(counters, (size_t)ld);
    DH_HIP(hipMemset(M, 0, sizeof(double) * (size_t)ld * ld));
    DH_HIP(hipMemset(x, 0, sizeof(cplx) * (size_t)ld * R));
    DH_HIP(hipMemset(x2, 0, sizeof(cplx) * (size_t)ld * R));
    DH_HIP(hipMemset(acc, 0, sizeof(cplx) * (size_t)ld * R));
    DH_HIP(hipMemset(coef, 0, sizeof(cplx) * 8));
    DH_HIP(hipMemset(counters, 0, sizeof(unsigned) * ld));
    MatvecArgs a{};
    a.M = M; a.ld = ld; a.N = N; a.R = R; a.x = x; a.x2 = x2; a.out = out; a.acc = acc; a.partial = partial;
    a.counters = counters; a.split = split; a.tri = TRI_FULL; a.epi = EPI_CHEB; a.alpha = 1.0; a.beta = 0.0;
    a.coef = coef; a.k = 2; a.x_stride = R; a.x_off = 0; a.out_stride = R; a.out_off = 0;
    hipEvent_t e0, e1;
    DH_HIP(hipEventCreate(&e0));
    DH_HIP(hipEventCreate(&e1));
    float* M32 = nullptr;
    if (variant < 0) {  // -1: FP32 matrix-core GEMM (mixed); -2: FP32-stored H, FP64 products
        DALLOC(M32, (size_t)ld * ld);
        DH_HIP(hipMemset(M32, 0, sizeof(float) * (size_t)ld * ld));
        a.M32 = M32;
    }
    auto run = [&]() {
        if (variant == -1) launch_gemm3(a, 1, reinterpret_cast<double*>(partial), nullptr);
        else if (variant == -2) launch_term_f32h(a, nullptr);
        else launch_matvec_variant(a, variant, nullptr);
    };
    for (int w = 0; w < 3; ++w) run();
    DH_HIP(hipEventRecord(e0, nullptr));
    for (int r = 0; r < reps; ++r) run();
    DH_HIP(hipEventRecord(e1, nullptr));
    DH_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    DH_HIP(hipEventElapsedTime(&ms, e0, e1));
    *us_per_launch = 1e3 * ms / reps;
    hipEventDestroy(e0); hipEventDestroy(e1);
    hipFree(M); hipFree(x); hipFree(x2); hipFree(out); hipFree(acc); hipFree(coef); hipFree(partial); hipFree(counters);
    if (M32) hipFree(M32);
    return DH_OK;
}

extern "C" int dh_set_precision(dh_system* s, int mode) {
    CHECK_ARG(s && mode >= 0 && mode <= 3,
              "dh_set_precision: mode must be 0 (FP64), 1 (mixed), 2 (FP32-stored operator) or 3 (mixed with the "
              "residual correction)");
    DH_HIP(hipSetDevice(s->ctx->device));
    if (mode != 0 && !s->H32) DALLOC(s->H32, (size_t)s->ld * s->ld);
    if (mode == 3 && !s->Hlo) DALLOC(s->Hlo, (size_t)s->ld * s->ld);
    // the FP32 and bf16 term inputs: rows N..ld-1 stay zero (the epilogues write
    // rows < N only, and the library GEMMs run over all ld rows)
    if (mode == 1 || mode == 3)
        for (float*& p : s->x32)
            if (!p) {
                DALLOC(p, (size_t)s->ld * 4 * s->W);
                launch_zero_cplx(reinterpret_cast<cplx*>(p), (long)s->ld * s->W, s->ctx->stream);
            }
    if (mode == 3)
        for (unsigned short*& p : s->xb16)
            if (!p) {
                DALLOC(p, (size_t)s->ld * 4 * s->W);
                launch_zero_cplx(reinterpret_cast<cplx*>(p), (long)s->ld * s->W / 2, s->ctx->stream);
            }
    // the FP32 products of modes 1 and 3 run on rocBLAS through the context's handle
    if ((mode == 1 || mode == 3) && !s->ctx->lib && !(s->ctx->lib = lib_create(s->ctx->stream))) {
        dh::set_error("dh_set_precision: rocBLAS handle creation failed");
        return DH_ERR_HIP;
    }
    if (mode != s->precision) s->h32_ready = false;  // mode 3 also needs the residual
    s->precision = mode;
    if (mode != 0 && s->ops_ready && !s->h32_ready) {
        if (mode == 3) launch_split_f32_bf16(s->Hth, s->H32, s->Hlo, (long)s->ld * s->ld, s->ctx->stream);
        else launch_to_f32(s->Hth, s->H32, (long)s->ld * s->ld, s->ctx->stream);
        DH_HIP(hipStreamSynchronize(s->ctx->stream));
        s->h32_ready = true;
    }
    return DH_OK;
}

// ------------------------------------------------------------------ multi-system batches
// Many systems (uneven N, same wavepacket count) advance together: one grouped
// launch per Chebyshev term covers every system that still has terms left
// (systems sorted by term count, so the active set is a prefix of the task
// table), then grouped norm checks, commits and observables.
struct dh_batch {
    dh_context* ctx = nullptr;
    std::vector<dh_system*> sys;
    int W = 0, R = 0, n = 0;
    // descriptor kinds: terms 1, 2, odd and even k >= 3 of a one-step series,
    // the outputs, and the terms k >= 3 of a multi-step series by k % PANEL
    // (ring slots)
    enum { D_K1 = 0, D_K2, D_KODD, D_KEVEN, D_ENERGY, D_LBRA, D_LINVT, D_RING0, NDESC = D_RING0 + PANEL };
    MatvecArgs* d_args = nullptr;  // [NDESC][n]
    // panel passes of a multi-step series: task (member, 256-element block),
    // members in term-count order (prefixP[j]: tasks of the first j)
    PanelArgs* d_panel = nullptr;  // [n]
    int *d_p_desc = nullptr, *d_p_blk = nullptr;
    std::vector<int> prefixP;
    std::vector<PanelArgs> h_panel;
    int *d_task_desc = nullptr, *d_task_blk = nullptr;
    NormTask* d_norms = nullptr;   // [2][n*R]: before, after
    CopyTask* d_copy = nullptr;    // [n]
    ObsTask* d_obs = nullptr;      // [n*W]
    double* d_nval = nullptr;      // [2][n*R]
    std::vector<int> order, prefix;  // systems by K descending; prefix task counts
    std::vector<int> Ks;
    int Kmax = 0;
    double stats[4] = {0, 0, 0, 0};  // steps, term launches, last max drift, retries
    // the term series as one captured graph, kept while its key (term counts,
    // kernel family, and the by-value arguments of the members that run their
    // own series) stays the same; the grouped kernels read the per-step
    // descriptors from d_args
    bool use_graph = true;
    hipGraphExec_t gexec = nullptr;
    std::vector<double> gkey;
    // grouped term kernel: symmetric half-storage (R <= 8; phase-1 tasks:
    // segments of up to 8 tiles, phase-2 tasks: block x chunk; both shapes a
    // function of the system's ld alone) or dense column blocks, as the batch
    // autotuner chose for this R
    bool sym_ok = false, sym = false;
    int *d_t1_desc = nullptr, *d_t1_idx = nullptr, *d_t2_desc = nullptr, *d_t2_idx = nullptr, *d_sysS = nullptr;
    std::vector<int> prefix1, prefix2;
    int reduce_groups = 4;  // phase-2 groups: those of the members' own launches (the largest)
    // members in a precision mode other than FP64 run their own term series
    // (FP32 matrix-core GEMM or FP32-stored operator) inside the batch step
    std::vector<char> solo;
    // host copies of the last uploaded tables: uploaded again only on change
    std::vector<MatvecArgs> h_args;
    std::vector<NormTask> h_norms;
    std::vector<CopyTask> h_copy;
    std::vector<ObsTask> h_obs;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

// Batch term kernel per R, chosen once per process by timing both grouped
// families on the first batch of that R (-1: not yet tuned, 0: dense column
// blocks, 1: symmetric half storage); every later batch of that R runs the
// same kernels, so a system's arithmetic does not depend on its batch.
static int g_batch_family[6] = {-1, -1, -1, -1, -1, -1};
static std::mutex g_batch_family_mu;
static int batch_family_slot(int R) { return R == 2 ? 1 : R == 4 ? 2 : R == 8 ? 3 : R == 16 ? 4 : R == 32 ? 5 : 0; }

static void batch_term_launch(dh_batch* b, bool sym, int kind, int active, int k, hipStream_t st) {
    const int n = b->n;
    if (sym)
        launch_sym_grouped(b->d_args + (size_t)kind * n, b->d_t1_desc, b->d_t1_idx, b->prefix1[active], b->d_t2_desc,
                           b->d_t2_idx, b->prefix2[active], b->d_sysS, b->R, k, b->reduce_groups, st);
    else
        launch_matvec_grouped(b->d_args + (size_t)kind * n, b->d_task_desc, b->d_task_blk, b->prefix[active], b->R, k,
                              st);
}

// Chebyshev terms of a batch's grouped members: those with K_s >= k are the
// first `active` of `order`.  A multi-step series (multi) writes its terms to
// the members' rings and adds each PANEL of them to the accumulators of the
// members that took any (every member's last panel ends at its own K_s).
static void batch_series_launches(dh_batch* b, hipStream_t st, bool multi = false) {
    const int n = b->n;
    int active = n, k_taken = -1;
    for (int k = 1; k <= b->Kmax; ++k) {
        while (active > 0 && b->Ks[b->order[active - 1]] < k) --active;
        const int kind = k == 1   ? dh_batch::D_K1
                         : k == 2 ? dh_batch::D_K2
                         : multi  ? dh_batch::D_RING0 + k % PANEL
                         : (k & 1) ? dh_batch::D_KODD
                                   : dh_batch::D_KEVEN;
        batch_term_launch(b, b->sym, kind, active, k, st);
        if (multi && (k % PANEL == 0 || k == b->Kmax)) {
            const int k0 = k_taken + 1;
            int took = n;
            while (took > 0 && b->Ks[b->order[took - 1]] < k0) --took;
            launch_multi_panel_grouped(b->d_panel, b->d_p_desc, b->d_p_blk, b->prefixP[took], k0, k, st);
            k_taken = k;
        }
    }
}

static std::vector<double> batch_graph_key(const dh_batch* b, bool multi) {
    std::vector<double> key(b->Ks.begin(), b->Ks.end());
    key.push_back(b->sym ? 1.0 : 0.0);
    key.push_back(multi ? 1.0 : 0.0);
    for (int i = 0; i < b->n; ++i) key.push_back(b->solo[i] ? 1.0 : 0.0);
    return key;
}

// members in other precision modes run their own series after the grouped
// terms, outside the capture (the FP32-product modes call rocBLAS)
static int batch_solo_series(dh_batch* b) {
    for (int i = 0; i < b->n; ++i)
        if (b->solo[i])
            if (int rc = series_launches(b->sys[i])) return rc;
    return DH_OK;
}

static int batch_series(dh_batch* b, hipStream_t st, bool multi = false) {
    if (!b->use_graph || b->Kmax == 0) {
        batch_series_launches(b, st, multi);
        return batch_solo_series(b);
    }
    std::vector<double> key = batch_graph_key(b, multi);
    if (!b->gexec || b->gkey != key) {
        if (b->gexec) {
            DH_HIP(hipGraphExecDestroy(b->gexec));
            b->gexec = nullptr;
        }
        DH_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        batch_series_launches(b, st, multi);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &graph);
        if (e != hipSuccess) {
            if (graph) hipGraphDestroy(graph);
            dh::set_error("batch term series capture failed: %s", hipGetErrorString(e));
            return DH_ERR_HIP;
        }
        const hipError_t ei = hipGraphInstantiate(&b->gexec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            b->gexec = nullptr;
            dh::set_error("batch term series graph instantiation failed: %s", hipGetErrorString(ei));
            return DH_ERR_HIP;
        }
        b->gkey = key;
    }
    DH_HIP(hipGraphLaunch(b->gexec, st));
    return batch_solo_series(b);
}

extern "C" int dh_batch_set_graph(dh_batch* b, int on) {
    CHECK_ARG(b, "dh_batch_set_graph: null batch");
    b->use_graph = on != 0;
    return DH_OK;
}

extern "C" int dh_batch_destroy(dh_batch* b) {
    if (!b) return DH_OK;
    hipSetDevice(b->ctx->device);
    hipStreamSynchronize(b->ctx->stream);
    if (b->gexec) hipGraphExecDestroy(b->gexec);
    if (b->ev0) hipEventDestroy(b->ev0);
    if (b->ev1) hipEventDestroy(b->ev1);
    void* ptrs[] = {b->d_args,  b->d_task_desc, b->d_task_blk, b->d_norms,  b->d_copy,  b->d_obs,  b->d_nval,
                    b->d_t1_desc, b->d_t1_idx,  b->d_t2_desc,  b->d_t2_idx, b->d_sysS, b->d_panel, b->d_p_desc,
                    b->d_p_blk};
    for (void* p : ptrs)
        if (p) hipFree(p);
    delete b;
    return DH_OK;
}

extern "C" int dh_batch_create(dh_context* ctx, int n, dh_system* const* systems, dh_batch** out) {
    CHECK_ARG(ctx && out && systems && n > 0, "dh_batch_create: bad argument");
    const int W = systems[0]->W;
    for (int i = 0; i < n; ++i) {
        CHECK_ARG(systems[i] && systems[i]->ctx == ctx, "dh_batch_create: systems must belong to the context");
        CHECK_ARG(systems[i]->W == W, "dh_batch_create: all systems need the same wavepacket count");
    }
    const int R = 2 * W;
    CHECK_ARG(R == 2 || R == 4 || R == 8 || R == 16 || R == 32,
              "dh_batch_create: wavepackets per system must be 1, 2, 4, 8 or 16");
    DH_HIP(hipSetDevice(ctx->device));
    dh_batch* b = new dh_batch();
    b->ctx = ctx;
    b->sys.assign(systems, systems + n);
    b->W = W;
    b->R = R;
    b->n = n;
    b->solo.assign(n, 0);
    long tasks = 0;
    const int cpb0 = std::min(grouped_cols_per_block(R), grouped_cols_per_block(W));
    for (auto* s : b->sys) tasks += (s->N + cpb0 - 1) / cpb0;
    int rc;
#define BALLOC(p, c)                                    \
    do {                                                \
        if ((rc = dalloc(&(p), (c))) != DH_OK) {        \
            dh_batch_destroy(b);                        \
            return rc;                                  \
        }                                               \
    } while (0)
    // term tasks (grouped members only) and output tasks (every member)
    BALLOC(b->d_args, (size_t)dh_batch::NDESC * n);
    BALLOC(b->d_task_desc, (size_t)2 * tasks);
    BALLOC(b->d_task_blk, (size_t)2 * tasks);
    BALLOC(b->d_norms, (size_t)2 * n * R);
    BALLOC(b->d_copy, (size_t)n);
    BALLOC(b->d_obs, (size_t)n * W);
    BALLOC(b->d_nval, (size_t)2 * n * R);
    long ptasks = 0;
    for (auto* s : b->sys) ptasks += ((long)s->N * R + 255) / 256;
    BALLOC(b->d_panel, (size_t)n);
    BALLOC(b->d_p_desc, (size_t)ptasks);
    BALLOC(b->d_p_blk, (size_t)ptasks);
    b->sym_ok = sym_supported(R, TRI_FULL);
    if (b->sym_ok) {
        long t1 = 0, t2 = 0;
        std::vector<int> sS(n);
        for (int i = 0; i < n; ++i) {
            t1 += sym_term_segments(b->sys[i]->ld, &sS[i]);
            t2 += (long)(b->sys[i]->ld / TILE) * R;
            b->reduce_groups = std::max(b->reduce_groups, sym_reduce_groups_for(b->sys[i]->ld));
        }
        BALLOC(b->d_t1_desc, (size_t)t1);
        BALLOC(b->d_t1_idx, (size_t)t1);
        BALLOC(b->d_t2_desc, (size_t)t2);
        BALLOC(b->d_t2_idx, (size_t)t2);
        BALLOC(b->d_sysS, (size_t)n);
        if (hipMemcpyAsync(b->d_sysS, sS.data(), sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess) {
            dh_batch_destroy(b);
            dh::set_error("dh_batch_create: copy of the segment table failed");
            return DH_ERR_HIP;
        }
    }
#undef BALLOC
    if (hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
        dh_batch_destroy(b);
        dh::set_error("dh_batch_create: event creation failed");
        return DH_ERR_HIP;
    }
    *out = b;
    return DH_OK;
}

static MatvecArgs batch_desc(dh_system* s, int kind) {
    const int W = s->W, R = 2 * W;
    MatvecArgs a{};
    a.ld = s->ld;
    a.N = s->N;
    a.coef = s->coef;
    a.acc = s->acc;
    a.partial = s->partial;  // the symmetric term's partial sums
    const double c = s->coef_c, r = s->coef_r;
    switch (kind) {
        case dh_batch::D_K1:
            a.M = s->Hth; a.R = R; a.tri = TRI_FULL; a.epi = EPI_CHEB_FIRST;
            a.x = s->phi; a.out = s->x1; a.alpha = 1.0 / r; a.beta = -c / r;
            a.x_stride = R; a.out_stride = R;
            break;
        case dh_batch::D_K2:
            a.M = s->Hth; a.R = R; a.tri = TRI_FULL; a.epi = EPI_CHEB;
            a.x = s->x1; a.x2 = s->phi; a.out = s->x0; a.alpha = 2.0 / r; a.beta = -2.0 * c / r;
            a.x_stride = R; a.out_stride = R;
            break;
        case dh_batch::D_KODD:  // k = 3, 5, ...: T_{k-1} in x0, T_{k-2} in x1, T_k over x1
            a.M = s->Hth; a.R = R; a.tri = TRI_FULL; a.epi = EPI_CHEB;
            a.x = s->x0; a.x2 = s->x1; a.out = s->x1; a.alpha = 2.0 / r; a.beta = -2.0 * c / r;
            a.x_stride = R; a.out_stride = R;
            break;
        case dh_batch::D_KEVEN:  // k = 4, 6, ...
            a.M = s->Hth; a.R = R; a.tri = TRI_FULL; a.epi = EPI_CHEB;
            a.x = s->x1; a.x2 = s->x0; a.out = s->x0; a.alpha = 2.0 / r; a.beta = -2.0 * c / r;
            a.x_stride = R; a.out_stride = R;
            break;
        case dh_batch::D_ENERGY:
            a.M = s->Hth; a.R = W; a.tri = TRI_FULL; a.epi = EPI_PLAIN;
            a.x = s->phi; a.x_stride = R; a.x_off = W; a.out = s->hk; a.out_stride = W;
            break;
        case dh_batch::D_LBRA:
            a.M = s->Q1; a.R = W; a.tri = TRI_UPPER; a.epi = EPI_PLAIN;
            a.x = s->phi; a.x_stride = R; a.out = s->out; a.out_stride = 3 * W;
            break;
        case dh_batch::D_LINVT:
            a.M = s->Q2; a.R = R; a.tri = TRI_LOWER; a.epi = EPI_PLAIN;
            a.x = s->phi; a.x_stride = R; a.out = s->out; a.out_stride = 3 * W; a.out_off = W;
            break;
        default:  // D_RING0 + k % PANEL, k >= 3 (set up by dh_batch_propagate_steps)
            a.M = s->Hth; a.R = R; a.tri = TRI_FULL; a.epi = EPI_CHEB;
            a.alpha = 2.0 / r; a.beta = -2.0 * c / r;
            a.x_stride = R; a.out_stride = R;
            break;
    }
    return a;
}

// a multi-step series' term descriptors (kinds D_K1, D_K2 and the ring kinds):
// T_k in ring slot (k - 1) % PANEL, no single accumulator
static MatvecArgs batch_ring_desc(dh_system* s, int kind) {
    const int R = 2 * s->W;
    MatvecArgs a = batch_desc(s, kind);
    const double c = 0.5 * (s->emax + s->emin), r = 0.5 * (s->emax - s->emin) * s->widen;
    const double f = kind == dh_batch::D_K1 ? 1.0 : 2.0;
    a.alpha = f / r;
    a.beta = -f * c / r;
    a.acc = nullptr;
    a.coef = nullptr;
    auto slot = [&](int k) { return s->ring + (long)((k - 1) % PANEL) * s->ld * R; };
    if (kind == dh_batch::D_K1) {
        a.x = s->phi;
        a.out = slot(1);
    } else if (kind == dh_batch::D_K2) {
        a.x = slot(1);
        a.x2 = s->phi;
        a.out = slot(2);
    } else {
        const int j = kind - dh_batch::D_RING0;
        const int k = j >= 3 ? j : j + PANEL;  // a term k >= 3 with k % PANEL == j
        a.x = slot(k - 1);
        a.x2 = slot(k - 2);
        a.out = slot(k);
    }
    return a;
}

// host -> device table copy, only when the content changed since the last copy
// (from the batch's own copy, which outlives the asynchronous transfer)
template <class T>
static int upload_if_changed(std::vector<T>& cache, const std::vector<T>& v, T* dst, hipStream_t st) {
    if (cache.size() == v.size() && (v.empty() || std::memcmp(cache.data(), v.data(), sizeof(T) * v.size()) == 0))
        return DH_OK;
    DH_HIP(hipStreamSynchronize(st));  // the previous upload from the cache is done
    cache = v;
    if (!cache.empty()) DH_HIP(hipMemcpyAsync(dst, cache.data(), sizeof(T) * cache.size(), hipMemcpyHostToDevice, st));
    return DH_OK;
}

// Times the first grouped term (k = 1: reads phi, writes x1 and acc, which the
// series then overwrites) with both kernel families and keeps the faster for R.
static int batch_autotune(dh_batch* b, hipStream_t st) {
    const int slot = batch_family_slot(b->R);
    if (b->prefix[b->n] == 0) {  // every member runs its own series: nothing to time
        b->sym = false;
        return DH_OK;
    }
    {
        std::lock_guard<std::mutex> lk(g_batch_family_mu);
        if (g_batch_family[slot] >= 0) {
            b->sym = g_batch_family[slot] == 1 && b->sym_ok;
            return DH_OK;
        }
    }
    int pick = 0;
    if (b->sym_ok) {
        float ms[2] = {0.f, 0.f};
        for (int fam = 0; fam < 2; ++fam) {
            for (int w = 0; w < 2; ++w) batch_term_launch(b, fam == 1, dh_batch::D_K1, b->n, 1, st);
            DH_HIP(hipEventRecord(b->ev0, st));
            for (int r = 0; r < 5; ++r) batch_term_launch(b, fam == 1, dh_batch::D_K1, b->n, 1, st);
            DH_HIP(hipEventRecord(b->ev1, st));
            DH_HIP(hipEventSynchronize(b->ev1));
            DH_HIP(hipEventElapsedTime(&ms[fam], b->ev0, b->ev1));
        }
        pick = ms[1] < ms[0] ? 1 : 0;
    }
    std::lock_guard<std::mutex> lk(g_batch_family_mu);
    if (g_batch_family[slot] < 0) g_batch_family[slot] = pick;
    b->sym = g_batch_family[slot] == 1 && b->sym_ok;
    return DH_OK;
}

extern "C" int dh_test_force_batch_family(int R, int family) {
    CHECK_ARG((R == 2 || R == 4 || R == 8 || R == 16 || R == 32) && family >= -1 && family <= 1,
              "dh_test_force_batch_family: bad argument");
    std::lock_guard<std::mutex> lk(g_batch_family_mu);
    g_batch_family[batch_family_slot(R)] = family;
    return DH_OK;
}

extern "C" int dh_batch_tuning(dh_batch* b, int* family) {
    CHECK_ARG(b && family, "dh_batch_tuning: null argument");
    *family = b->sym ? 1 : 0;
    return DH_OK;
}

// Task tables of a batch for its members' term counts Ks (descending order;
// members that run their own series add no grouped term tasks)
static int batch_retable(dh_batch* b, const std::vector<int>& Ks, const std::vector<char>& solo, hipStream_t st) {
    const int n = b->n, R = b->R, W = b->W;
    b->Ks = Ks;
    b->solo = solo;
    b->order.resize(n);
    for (int i = 0; i < n; ++i) b->order[i] = i;
    std::stable_sort(b->order.begin(), b->order.end(), [&](int x, int y) { return Ks[x] > Ks[y]; });
    std::vector<int> td, tb;
    b->prefix.assign(1, 0);
    const int cpb = std::min(grouped_cols_per_block(R), grouped_cols_per_block(W));
    for (int q : b->order) {
        const int nbq = solo[q] ? 0 : (b->sys[q]->N + cpb - 1) / cpb;
        for (int blk = 0; blk < nbq; ++blk) {
            td.push_back(q);
            tb.push_back(blk);
        }
        b->prefix.push_back((int)td.size());
    }
    // output tasks: every member, after the term tasks
    const int nterm = (int)td.size();
    for (int q = 0; q < n; ++q)
        for (int blk = 0; blk < (b->sys[q]->N + cpb - 1) / cpb; ++blk) {
            td.push_back(q);
            tb.push_back(blk);
        }
    b->prefix.push_back(nterm);  // prefix[n + 1]: start of the output tasks
    b->prefix.push_back((int)td.size());
    DH_HIP(hipMemcpyAsync(b->d_task_desc, td.data(), sizeof(int) * td.size(), hipMemcpyHostToDevice, st));
    DH_HIP(hipMemcpyAsync(b->d_task_blk, tb.data(), sizeof(int) * tb.size(), hipMemcpyHostToDevice, st));
    if (b->sym_ok) {  // the same member order for the symmetric term's two phases
        std::vector<int> d1, i1, d2, i2;
        b->prefix1.assign(1, 0);
        b->prefix2.assign(1, 0);
        for (int q : b->order) {
            if (!solo[q]) {
                int S = 0;
                const int nseg = sym_term_segments(b->sys[q]->ld, &S);
                for (int g = 0; g < nseg; ++g) {
                    d1.push_back(q);
                    i1.push_back(g);
                }
                const int nbq = b->sys[q]->ld / TILE;
                for (int v = 0; v < nbq * R; ++v) {
                    d2.push_back(q);
                    i2.push_back(v);
                }
            }
            b->prefix1.push_back((int)d1.size());
            b->prefix2.push_back((int)d2.size());
        }
        if (!d1.empty()) {
            DH_HIP(hipMemcpyAsync(b->d_t1_desc, d1.data(), sizeof(int) * d1.size(), hipMemcpyHostToDevice, st));
            DH_HIP(hipMemcpyAsync(b->d_t1_idx, i1.data(), sizeof(int) * i1.size(), hipMemcpyHostToDevice, st));
            DH_HIP(hipMemcpyAsync(b->d_t2_desc, d2.data(), sizeof(int) * d2.size(), hipMemcpyHostToDevice, st));
            DH_HIP(hipMemcpyAsync(b->d_t2_idx, i2.data(), sizeof(int) * i2.size(), hipMemcpyHostToDevice, st));
        }
    }
    {  // panel passes: every grouped member, 256 elements of rows < N per task
        std::vector<int> pd, pb;
        b->prefixP.assign(1, 0);
        for (int q : b->order) {
            if (!solo[q])
                for (long blk = 0; blk < ((long)b->sys[q]->N * R + 255) / 256; ++blk) {
                    pd.push_back(q);
                    pb.push_back((int)blk);
                }
            b->prefixP.push_back((int)pd.size());
        }
        if (!pd.empty()) {
            DH_HIP(hipMemcpyAsync(b->d_p_desc, pd.data(), sizeof(int) * pd.size(), hipMemcpyHostToDevice, st));
            DH_HIP(hipMemcpyAsync(b->d_p_blk, pb.data(), sizeof(int) * pb.size(), hipMemcpyHostToDevice, st));
        }
        DH_HIP(hipStreamSynchronize(st));  // pd, pb are local
    }
    b->Kmax = 0;
    for (int i = 0; i < n; ++i)
        if (!solo[i]) b->Kmax = std::max(b->Kmax, Ks[i]);
    DH_HIP(hipStreamSynchronize(st));  // the task vectors above are pageable and local
    return DH_OK;
}

extern "C" int dh_batch_propagate(dh_batch* b, double t_init, double t_max, int* max_terms) {
    CHECK_ARG(b, "dh_batch_propagate: null batch");
    DH_HIP(hipSetDevice(b->ctx->device));
    hipStream_t st = b->ctx->stream;
    const double dt = (t_max - t_init) / H_BAR;
    const int n = b->n, R = b->R, W = b->W;
    for (auto* s : b->sys) {
        if (!s->ops_ready || !s->state_set) {
            dh::set_error("dh_batch_propagate: every system needs operators and wavepackets");
            return DH_ERR_STATE;
        }
    }
    if (!(dt > 0.0)) return DH_OK;
    std::vector<char> solo(n);
    for (int i = 0; i < n; ++i) {
        dh_system* s = b->sys[i];
        solo[i] = s->precision != 0;
        if (s->phi_dirty) {
            convert_ao_to_phi(s);
            s->phi_dirty = false;
        }
        int rc = ensure_coefficients(s, dt);
        if (rc) return rc;
    }
    // task tables by term count (descending); members in other precision modes
    // add no grouped term tasks
    std::vector<int> Ks(n);
    for (int i = 0; i < n; ++i) Ks[i] = b->sys[i]->K;
    const bool retable = Ks != b->Ks || solo != b->solo || b->prefix.empty();
    if (retable)
        if (int rc = batch_retable(b, Ks, solo, st)) return rc;
    // per-member descriptors, norm tasks: uploaded when they change (bounds,
    // buffers); pointers are read on the device every step
    std::vector<MatvecArgs> args((size_t)dh_batch::NDESC * n);
    for (int kind = 0; kind < dh_batch::NDESC; ++kind)
        for (int i = 0; i < n; ++i) args[(size_t)kind * n + i] = batch_desc(b->sys[i], kind);
    std::vector<NormTask> nt((size_t)2 * n * R);
    for (int i = 0; i < n; ++i)
        for (int c = 0; c < R; ++c) {
            nt[(size_t)i * R + c] = NormTask{b->sys[i]->phi, b->sys[i]->N, R, c, b->d_nval + (size_t)i * R + c};
            nt[(size_t)n * R + (size_t)i * R + c] =
                NormTask{b->sys[i]->acc, b->sys[i]->N, R, c, b->d_nval + (size_t)n * R + (size_t)i * R + c};
        }
    if (int rc = upload_if_changed(b->h_args, args, b->d_args, st)) return rc;
    if (int rc = upload_if_changed(b->h_norms, nt, b->d_norms, st)) return rc;
    if (retable || b->gkey.empty())
        if (int rc = batch_autotune(b, st)) return rc;
    launch_grouped_norms(b->d_norms, n * R, st);
    if (int rc = batch_series(b, st)) return rc;
    launch_grouped_norms(b->d_norms + (size_t)n * R, n * R, st);
    std::vector<double> nv((size_t)2 * n * R);
    DH_HIP(hipMemcpyAsync(nv.data(), b->d_nval, sizeof(double) * nv.size(), hipMemcpyDeviceToHost, st));
    DH_HIP(hipStreamSynchronize(st));
    std::vector<int> ok(n, 1);
    double maxdrift = 0.0;
    for (int i = 0; i < n; ++i) {
        const double tol = fp32_products(b->sys[i]) ? 1e-4 : 1e-9;
        for (int c = 0; c < R; ++c) {
            const double n0 = nv[(size_t)i * R + c], n1 = nv[(size_t)n * R + (size_t)i * R + c];
            const double d = n0 > 0 ? std::fabs(n1 - n0) / n0 : 0.0;
            if (!(d < tol)) ok[i] = 0;
            if (std::isfinite(d)) maxdrift = std::max(maxdrift, d);
        }
    }
    // commit acc -> phi for systems whose norm held
    std::vector<CopyTask> ct;
    for (int i = 0; i < n; ++i)
        if (ok[i]) ct.push_back(CopyTask{b->sys[i]->acc, b->sys[i]->phi, (long)b->sys[i]->N * R});
    if (int rc = upload_if_changed(b->h_copy, ct, b->d_copy, st)) return rc;
    launch_grouped_copy(b->d_copy, (int)ct.size(), st);
    // the mixed mode's per-step FP64 norm restoration, as in its single step
    for (int i = 0; i < n; ++i)
        if (ok[i] && fp32_products(b->sys[i]))
            launch_renorm_columns(b->sys[i]->phi, b->sys[i]->N, R, b->d_nval + (size_t)i * R,
                                  b->d_nval + (size_t)n * R + (size_t)i * R, st);
    // outputs of every member
    const int o0 = b->prefix[n + 1], o1 = b->prefix[n + 2];
    launch_matvec_grouped(b->d_args + (size_t)dh_batch::D_ENERGY * n, b->d_task_desc + o0, b->d_task_blk + o0, o1 - o0,
                          W, 0, st);
    launch_matvec_grouped(b->d_args + (size_t)dh_batch::D_LBRA * n, b->d_task_desc + o0, b->d_task_blk + o0, o1 - o0, W,
                          0, st);
    launch_matvec_grouped(b->d_args + (size_t)dh_batch::D_LINVT * n, b->d_task_desc + o0, b->d_task_blk + o0, o1 - o0,
                          R, 0, st);
    std::vector<ObsTask> ot((size_t)n * W);
    for (int i = 0; i < n; ++i) {
        dh_system* s = b->sys[i];
        for (int w = 0; w < W; ++w)
            ot[(size_t)i * W + w] = ObsTask{s->N, W, s->n_frag, w, s->frag, s->out, s->phi, s->hk, s->pops, s->erg, s->norms};
    }
    if (int rc = upload_if_changed(b->h_obs, ot, b->d_obs, st)) return rc;
    launch_grouped_observables(b->d_obs, n * W, st);
    DH_HIP(hipGetLastError());
    DH_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < n; ++i) {
        dh_system* s = b->sys[i];
        if (ok[i]) {
            s->stats[0] = s->K + 1;
            s->stats[1] += 1;
            s->stats[2] += s->K + 3;
            continue;
        }
        // bounds did not enclose this system's spectrum: redo it alone with widening
        s->widen *= 1.25;
        s->stats[3] += 1;
        b->stats[3] += 1;
        int rc = dh_propagate(s, t_init, t_max, nullptr);
        if (rc) return rc;
    }
    b->stats[0] += 1;
    b->stats[1] += b->Kmax;
    b->stats[2] = maxdrift;
    if (max_terms) *max_terms = *std::max_element(Ks.begin(), Ks.end()) + 1;
    return DH_OK;
}

// n_steps steps of dt_ps for every member, up to MULTI_MAX steps per
// recurrence (the multi-step series of dh_propagate_steps, grouped): the
// grouped terms write each member's ring, and grouped panel passes add them to
// the member's accumulators with its own coefficient rows and term table.
// One norm check, commit and output pass per recurrence; the outputs are those
// of the final state.  Batches whose members are not all FP64 step one step at
// a time (dh_batch_propagate).
extern "C" int dh_batch_propagate_steps(dh_batch* b, int n_steps, double dt_ps, int* max_terms) {
    CHECK_ARG(b && n_steps > 0 && dt_ps > 0.0, "dh_batch_propagate_steps: need a batch, n_steps > 0 and dt > 0");
    DH_HIP(hipSetDevice(b->ctx->device));
    hipStream_t st = b->ctx->stream;
    const double dt = dt_ps / H_BAR;
    const int n = b->n, R = b->R;
    int done = 0, kmax_seen = 0;
    bool all_fp64 = true;
    for (auto* s : b->sys) all_fp64 = all_fp64 && s->precision == 0;
    // the first step tunes the batch (kernel family, task tables) as usual
    if (b->prefix.empty() || !all_fp64 || !b->sym_ok) {
        int k = 0;
        if (int rc = dh_batch_propagate(b, 0.0, dt_ps, &k)) return rc;
        kmax_seen = k;
        done = 1;
    }
    if (!all_fp64) {
        for (; done < n_steps; ++done) {
            int k = 0;
            if (int rc = dh_batch_propagate(b, 0.0, dt_ps, &k)) return rc;
            kmax_seen = std::max(kmax_seen, k);
        }
        if (max_terms) *max_terms = kmax_seen;
        return DH_OK;
    }
    const long cs = KMAX_ALLOC + 2;
    while (done < n_steps) {
        int m = std::min(MULTI_MAX, n_steps - done);
        for (auto* s : b->sys) {
            if (s->phi_dirty) {
                convert_ao_to_phi(s);
                s->phi_dirty = false;
            }
            m = std::min(m, multi_chunk_steps(s, dt, m));
        }
        if (m < 2) {  // a member's coefficient table holds one step only
            int k = 0;
            if (int rc = dh_batch_propagate(b, 0.0, dt_ps, &k)) return rc;
            kmax_seen = std::max(kmax_seen, k);
            ++done;
            continue;
        }
        std::vector<int> Ks(n);
        for (int i = 0; i < n; ++i) {
            int Kq[MULTI_MAX];
            if (int rc = multi_coefficients(b->sys[i], dt, m, Kq)) return rc;
            Ks[i] = Kq[m - 1];
        }
        std::vector<char> solo(n, 0);
        if (Ks != b->Ks || solo != b->solo || b->prefix.empty())
            if (int rc = batch_retable(b, Ks, solo, st)) return rc;
        // term descriptors over the members' rings; the output descriptors
        // read phi, which the commit below fills
        std::vector<MatvecArgs> args((size_t)dh_batch::NDESC * n);
        std::vector<PanelArgs> pargs(n);
        for (int kind = 0; kind < dh_batch::NDESC; ++kind)
            for (int i = 0; i < n; ++i) {
                dh_system* s = b->sys[i];
                const bool term = kind == dh_batch::D_K1 || kind == dh_batch::D_K2 || kind >= dh_batch::D_RING0;
                args[(size_t)kind * n + i] = term ? batch_ring_desc(s, kind) : batch_desc(s, kind);
            }
        for (int i = 0; i < n; ++i) {
            dh_system* s = b->sys[i];
            PanelArgs p{};
            p.accm = s->accm;
            p.accm_stride = (long)s->ld * R;
            p.coefm = s->coefm;
            p.coefm_stride = (int)cs;
            p.ring = s->ring;
            p.ring_stride = (long)s->ld * R;
            p.phi = s->phi;
            p.kq = s->kq_d;
            p.m = m;
            p.n = (long)s->N * R;
            pargs[i] = p;
        }
        if (int rc = upload_if_changed(b->h_panel, pargs, b->d_panel, st)) return rc;
        std::vector<NormTask> nt((size_t)2 * n * R);
        for (int i = 0; i < n; ++i) {
            dh_system* s = b->sys[i];
            const cplx* last = s->accm + (size_t)(m - 1) * s->ld * R;
            for (int c = 0; c < R; ++c) {
                nt[(size_t)i * R + c] = NormTask{s->phi, s->N, R, c, b->d_nval + (size_t)i * R + c};
                nt[(size_t)n * R + (size_t)i * R + c] = NormTask{last, s->N, R, c, b->d_nval + (size_t)n * R + (size_t)i * R + c};
            }
        }
        if (int rc = upload_if_changed(b->h_args, args, b->d_args, st)) return rc;
        if (int rc = upload_if_changed(b->h_norms, nt, b->d_norms, st)) return rc;
        launch_grouped_norms(b->d_norms, n * R, st);
        if (int rc = batch_series(b, st, true)) return rc;
        launch_grouped_norms(b->d_norms + (size_t)n * R, n * R, st);
        std::vector<double> nv((size_t)2 * n * R);
        DH_HIP(hipMemcpyAsync(nv.data(), b->d_nval, sizeof(double) * nv.size(), hipMemcpyDeviceToHost, st));
        DH_HIP(hipStreamSynchronize(st));
        std::vector<int> ok(n, 1);
        double maxdrift = 0.0;
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < R; ++c) {
                const double n0 = nv[(size_t)i * R + c], n1 = nv[(size_t)n * R + (size_t)i * R + c];
                const double d = n0 > 0 ? std::fabs(n1 - n0) / n0 : 0.0;
                if (!(d < 1e-9)) ok[i] = 0;
                if (std::isfinite(d)) maxdrift = std::max(maxdrift, d);
            }
        // commit the last step's state for members whose norm held
        std::vector<CopyTask> ct;
        for (int i = 0; i < n; ++i)
            if (ok[i]) {
                dh_system* s = b->sys[i];
                ct.push_back(CopyTask{s->accm + (size_t)(m - 1) * s->ld * R, s->phi, (long)s->N * R});
            }
        if (int rc = upload_if_changed(b->h_copy, ct, b->d_copy, st)) return rc;
        launch_grouped_copy(b->d_copy, (int)ct.size(), st);
        const int W = b->W;
        const int o0 = b->prefix[n + 1], o1 = b->prefix[n + 2];
        launch_matvec_grouped(b->d_args + (size_t)dh_batch::D_ENERGY * n, b->d_task_desc + o0, b->d_task_blk + o0,
                              o1 - o0, W, 0, st);
        launch_matvec_grouped(b->d_args + (size_t)dh_batch::D_LBRA * n, b->d_task_desc + o0, b->d_task_blk + o0,
                              o1 - o0, W, 0, st);
        launch_matvec_grouped(b->d_args + (size_t)dh_batch::D_LINVT * n, b->d_task_desc + o0, b->d_task_blk + o0,
                              o1 - o0, R, 0, st);
        std::vector<ObsTask> ot((size_t)n * W);
        for (int i = 0; i < n; ++i) {
            dh_system* s = b->sys[i];
            for (int w = 0; w < W; ++w)
                ot[(size_t)i * W + w] =
                    ObsTask{s->N, W, s->n_frag, w, s->frag, s->out, s->phi, s->hk, s->pops, s->erg, s->norms};
        }
        if (int rc = upload_if_changed(b->h_obs, ot, b->d_obs, st)) return rc;
        launch_grouped_observables(b->d_obs, n * W, st);
        DH_HIP(hipGetLastError());
        DH_HIP(hipStreamSynchronize(st));
        for (int i = 0; i < n; ++i) {
            dh_system* s = b->sys[i];
            if (ok[i]) {
                s->stats[0] = Ks[i] + 1;
                s->stats[1] += m;
                s->stats[2] += Ks[i] + 3;
                continue;
            }
            // bounds did not enclose this member's spectrum: its m steps alone,
            // with the checked step loop (wider bounds where needed)
            s->stats[3] += 1;
            b->stats[3] += 1;
            if (int rc = dh_propagate_steps(s, m, dt_ps, 0, nullptr, 0.0, 0.0, nullptr, nullptr, nullptr)) return rc;
        }
        b->stats[0] += m;
        b->stats[1] += b->Kmax;
        b->stats[2] = maxdrift;
        kmax_seen = std::max(kmax_seen, b->Kmax + 1);
        done += m;
    }
    if (max_terms) *max_terms = kmax_seen;
    return DH_OK;
}

// preprocess_ElHl_Chebyshev (ElHl_Chebyshev.f:110-129): Psi_ket = C_ket, Psi_bra = S C_bra,
// i.e. phi_ket = L^T C_ket and phi_bra = L^-1 S C_bra = L^T C_bra, then the t = 0 outputs.
// Diabatic-Ehrenfest kernel matrix from the resident state and H'
// (diabatic-Ehren.f:100-116, :366-392; ehrenfestkernel2_gpu_, Taylor_gpu.cpp:800-875).
// AO_bra = conj(S^-1 Psi_bra) and AO_ket = Psi_ket are the propagator's outputs
// (ElHl_Chebyshev.f:277-279); A is built from them on the device, and the
// context's W buffer (free once H' is formed) holds it.
extern "C" int dh_ehrenfest_kernel(dh_system* s, int wp_el, int wp_hl, const double* X, int ldx, double* K, int ldk) {
    CHECK_ARG(s, "dh_ehrenfest_kernel: null system");
    CHECK_ARG(wp_el >= 0 && wp_el < s->W && wp_hl >= -1 && wp_hl < s->W,
              "dh_ehrenfest_kernel: wavepacket index out of range");
    CHECK_ARG((!K || ldk >= s->N) && (!X || ldx >= s->N), "dh_ehrenfest_kernel: leading dimension < N");
    dh_context* c = s->ctx;
    if (!s->ops_ready || c->ws_owner != s || c->ws_ld < s->ld || !(s->hprime_ready || s->w_ready)) {
        dh::set_error("dh_ehrenfest_kernel: the operators of this system are not in the context workspace "
                      "(build or load them again)");
        return DH_ERR_STATE;
    }
    if (!s->state_set) {
        dh::set_error("dh_ehrenfest_kernel: no wavepackets set");
        return DH_ERR_STATE;
    }
    if (!X && s->n_atoms <= 0) {  // IP, k_WH and V_shift are filled only for systems with a basis
        dh::set_error("dh_ehrenfest_kernel: X not given and the system has no basis to form it from");
        return DH_ERR_STATE;
    }
    DH_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    const int W = s->W, ld = s->ld;
    if (!c->Kd) DALLOC(c->Kd, (size_t)c->ws_ld * c->ws_ld);
    if (!s->hprime_ready) {  // W = L^-1 h is still there; Q2 is mirrored by now
        launch_copy_lower(s->Q2, c->Kd, ld, st);
        launch_hprime(s, c->Kd);
    }
    s->w_ready = false;  // W holds A from here on
    if (s->phi_dirty) {  // states set in the AO basis: the outputs follow from phi
        convert_ao_to_phi(s);
        s->phi_dirty = false;
        launch_outputs(s);
    }
    if (X) {  // the caller's X goes into K, which the elementwise pass overwrites in place with X o A
        int rc = copy_matrix(c->Kd, ld, X, ldx, s->N, hipMemcpyHostToDevice, st);
        if (rc) return rc;
    }
    EhrenfestArgs e{};
    e.N = s->N; e.ld = ld;
    e.bra = s->out; e.bra_stride = 3 * W; e.bra_el = W + wp_el; e.bra_hl = wp_hl >= 0 ? W + wp_hl : -1;
    e.conj_bra = 1;
    e.ket = s->out; e.ket_stride = 3 * W; e.ket_el = 2 * W + wp_el; e.ket_hl = wp_hl >= 0 ? 2 * W + wp_hl : -1;
    e.X = X ? c->Kd : nullptr; e.ldx = ld;
    e.ip = s->bd.ip; e.kwh = s->bd.kwh; e.vshift = s->bd.vshift;
    e.A = c->W; e.K = c->Kd;
    if (!c->lib && !(c->lib = lib_create(st))) {
        dh::set_error("dh_ehrenfest_kernel: rocBLAS handle creation failed");
        return DH_ERR_HIP;
    }
    if (int rc = launch_ehrenfest(e, c->Hp, c->lib, st)) return rc;
    DH_HIP(hipGetLastError());
    if (K) {
        int rc = copy_matrix(K, ldk, c->Kd, ld, s->N, hipMemcpyDeviceToHost, st);
        if (rc) return rc;
    }
    DH_HIP(hipStreamSynchronize(st));
    return DH_OK;
}

// Eigenpairs il..iu (1-based, ascending) of the pencil (h, S) of the system's
// current operators: the generalised eigenproblem FMO solves for the initial
// electron and hole orbitals (FMO.f:294-350; xPU_dsygvd, GPU_Interface.cpp:705-853).
extern "C" int dh_eigenpairs(dh_system* s, int il, int iu, double* E, double* C, int ldc) {
    CHECK_ARG(s && il >= 1 && il <= iu && iu <= s->N, "dh_eigenpairs: need 1 <= il <= iu <= N");
    CHECK_ARG(!C || ldc >= s->N, "dh_eigenpairs: ldc < N");
    if (!s->ops_ready) {
        dh::set_error("dh_eigenpairs: operators not built");
        return DH_ERR_STATE;
    }
    dh_context* c = s->ctx;
    DH_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    const int ld = s->ld, N = s->N, nev = iu - il + 1;
    int rc = ensure_ws(c, ld);
    if (rc) return rc;
    if (!c->Kd) DALLOC(c->Kd, (size_t)c->ws_ld * c->ws_ld);
    if (!c->lib && !(c->lib = lib_create(st))) {
        dh::set_error("dh_eigenpairs: rocBLAS handle creation failed");
        return DH_ERR_HIP;
    }
    double *Z = nullptr, *Cd = nullptr, *Ed = nullptr;
    int* ints = nullptr;
    auto release = [&]() {
        hipFree(Z); hipFree(Cd); hipFree(Ed); hipFree(ints);
    };
    if (dalloc(&Z, (size_t)ld * ld) || dalloc(&Cd, (size_t)ld * nev) || dalloc(&Ed, 2 * (size_t)N) || dalloc(&ints, 2)) {
        release();
        return DH_ERR_NOMEM;
    }
    rc = eigen_pairs(c->lib, N, ld, s->Hth, s->Q2, c->Kd, Z, Cd, Ed, il, iu, ints, st);
    if (rc == DH_OK && E)
        if (hipMemcpyAsync(E, Ed + (il - 1), sizeof(double) * nev, hipMemcpyDeviceToHost, st) != hipSuccess)
            rc = DH_ERR_HIP;
    if (rc == DH_OK && C) rc = copy_block(C, ldc, Cd, ld, N, nev, hipMemcpyDeviceToHost, st) ? DH_ERR_HIP : DH_OK;
    if (hipStreamSynchronize(st) != hipSuccess && rc == DH_OK) rc = DH_ERR_HIP;
    release();
    if (rc == DH_ERR_HIP && !*dh::last_error()) dh::set_error("dh_eigenpairs: device copy failed");
    return rc;
}

// Wavepacket wp <- eigenvector k (1-based) of (h, S), as FMO seeds the electron
// and the hole with the whole-system LUMO and HOMO (tuning.f:76-83; FMO.f:294-350),
// then preprocess_ElHl_Chebyshev sets AO_bra = AO_ket = C (ElHl_Chebyshev.f:110-129).
extern "C" int dh_set_eigenstate(dh_system* s, int wp, int k, double* energy) {
    CHECK_ARG(s && wp >= 0 && wp < s->W && k >= 1 && k <= s->N, "dh_set_eigenstate: bad argument");
    std::vector<double> cr(s->N), cz(2 * (size_t)s->N, 0.0);
    double e = 0.0;
    int rc = dh_eigenpairs(s, k, k, &e, cr.data(), s->N);
    if (rc) return rc;
    for (int i = 0; i < s->N; ++i) cz[2 * (size_t)i] = cr[i];
    if (energy) *energy = e;
    return dh_set_initial_state(s, wp, cz.data(), cz.data());
}

extern "C" int dh_set_initial_state(dh_system* s, int wp, const double* ao_bra, const double* ao_ket) {
    CHECK_ARG(s && ao_bra && ao_ket && wp >= 0 && wp < s->W, "dh_set_initial_state: bad argument");
    CHECK_ARG(s->ops_ready, "dh_set_initial_state: operators not built");
    DH_HIP(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    const int W = s->W, R = 2 * W;
    if (s->phi_dirty && s->state_set) {  // bring the other wavepackets into the current basis first
        convert_ao_to_phi(s);
    }
    s->phi_dirty = false;
    // stage C_bra, C_ket in the hk scratch (N x W is enough for one column pair when W >= 2; use x1 otherwise)
    cplx* stage = s->x1;  // N x R scratch
    int prc = put_column(s, stage + 0, R, ao_bra, 0);
    if (prc) return prc;
    prc = put_column(s, stage + 1, R, ao_ket, 1);
    if (prc) return prc;
    MatvecArgs a{};
    a.ld = s->ld; a.N = s->N; a.partial = s->partial; a.counters = s->counters; a.epi = EPI_PLAIN;
    a.M = s->Q1; a.tri = TRI_LOWER; a.R = 1; a.split = matvec_split_for(s->N, 1);
    a.x = stage; a.x_stride = R; a.x_off = 0; a.out = s->phi; a.out_stride = R; a.out_off = wp;
    launch_matvec(a, st);
    a.x_off = 1; a.out_off = W + wp;
    launch_matvec(a, st);
    launch_outputs(s);
    DH_HIP(hipGetLastError());
    DH_HIP(hipStreamSynchronize(st));
    s->state_set = true;
    return DH_OK;
}
