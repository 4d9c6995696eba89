// Solver driver and C ABI of libqms.so.
//
// Per time step (QMS_3D.f90:1424-2234 restated matrix-free):
//   b  <- RHS(A^n, v)                     (k_rhs)
//   x  <- warm start with Dirichlet values (k_setbc)
//   FGMRES(m) on D^-1 M x = D^-1 b, restarted with an FP64 residual:
//     each iteration: z = P^-1 v (block-triangular multigrid preconditioner,
//     FP32 or FP64), w = D^-1 M z (fused FP64 operator), classical
//     Gram-Schmidt with fused multi-dot kernels.
// D = diag(M) equilibrates rows whose scales span 1e-3..1e15.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <string>
#include <vector>
#include <stdexcept>
#include <memory>
#include <functional>
#include <array>
#include <atomic>
#include <thread>
#include <algorithm>
#include <unordered_map>
#include <type_traits>
#include "qms_kernels.h"
#include "qms_plan.h"
#include "../../include/qms.h"

using namespace qms;

namespace {

struct HipError : std::runtime_error {
    int code;
    HipError(const std::string &m, int c) : std::runtime_error(m), code(c) {}
};

void check(hipError_t e, const char *what) {
    if (e == hipErrorOutOfMemory) throw HipError(std::string(what) + ": out of device memory", QMS_ERR_NOMEM);
    if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e), QMS_ERR_HIP);
}

thread_local std::string g_last_error;

// Mean milliseconds per call of f over `reps` calls on stream s, after one
// untimed call, with HIP events (released on every exit path).
double time_on(hipStream_t s, int reps, const std::function<void()> &f) {
    struct Events {
        hipEvent_t e[2] = {nullptr, nullptr};
        ~Events() {
            for (hipEvent_t x : e)
                if (x) hipEventDestroy(x);
        }
    } ev;
    check(hipEventCreate(&ev.e[0]), "event");
    check(hipEventCreate(&ev.e[1]), "event");
    f();
    check(hipEventRecord(ev.e[0], s), "event");
    for (int r = 0; r < reps; ++r) f();
    check(hipEventRecord(ev.e[1], s), "event");
    check(hipEventSynchronize(ev.e[1]), "event");
    float ms = 0;
    check(hipEventElapsedTime(&ms, ev.e[0], ev.e[1]), "event");
    return ms / reps;
}

}  // namespace

// The message qms_last_error(NULL) returns, for the C-ABI calls of other files.
void qms::set_last_error(const char *msg) { g_last_error = msg; }

// Krylov basis: V_0..V_m and the vector w being orthogonalised (the
// operator's output).  Each vector holds the A components as 3 n1 values of
// type TE from byte 0 and the V component as n1 FP64 values from byte offset
// offv: TE = double is a flat FP64 vector (offv = 24 n1); TE = float is the
// mixed-precision basis, whose V component stays FP64 because the
// conductors' floating potentials, fixed only through the 1e-3 S/m
// background, are below FP32 resolution of the conductor rows (an FP32 V
// component stalls the transient solves near a 1e-8 relative residual).
template <typename TE>
struct KBasis {
    std::vector<TE *> V;
    TE **dV = nullptr;
    TE *w = nullptr;
    int64_t offv = 0;
};

// Preconditioner state.  The A block's multigrid and the FGMRES Z basis
// (z_j = P^-1 v_j) are in TA (float in mixed precision, double otherwise);
// the V block's multigrid is always FP64: the conductors' floating
// potentials (a large constant on each conductor, fixed only through the
// 1e-3 S/m background) do not survive FP32 storage of the V correction, and
// FP32 there costs the Krylov solve about half as many iterations again.
template <typename TA>
struct PrecondState {
    Hierarchy<TA> hA[2];       // A block: steady (0) / transient (1) operator
    Hierarchy<double> hV;      // V block (independent of dt)
    bool builtA[2] = {false, false};
    bool builtV = false;
    TA *rT = nullptr;          // D v on the A rows, in TA
    double *rV = nullptr, *zV = nullptr;   // V-block right-hand side and correction
    std::vector<TA *> Z;
    TA **dZ = nullptr;
};

struct qms_handle {
    qms_config cfg;
    Grid g;
    BC bc;
    hipStream_t s = nullptr;
    std::vector<void *> allocs;
    int64_t bytes = 0;
    std::string err;

    uint16_t *slot = nullptr;
    double *mat = nullptr;
    Coeffs coef;
    CoeffsD c64;
    CoeffsF c32;
    std::vector<double> hcoef;  // host copy of the six FP64 coefficient arrays (multigrid setup)

    int64_t n = 0;              // 4*n1
    uint16_t *cls = nullptr;    // coefficient class per node (null: not built)
    double *ctab = nullptr;     // class table, 6 doubles per class
    int ncls = 0;
    bool use_cls = true;
    double *x = nullptr, *b = nullptr, *r = nullptr, *w = nullptr, *tmp = nullptr;
    double *Dst = nullptr, *Dtr = nullptr;   // row scaling, steady / transient
    double *xh1 = nullptr, *xh2 = nullptr;   // solutions of the last two steps (initial guess)
    int nhist = 0;                           // how many of them are valid
    // Krylov bases (allocated on first use): FP64 for the FP64 solve, FP32
    // for the mixed-precision inner solve
    KBasis<double> kb64;
    KBasis<float> kb32;
    double *partial = nullptr, *dsmall = nullptr;
    double *hsmall = nullptr;   // pinned
    // FGMRES bookkeeping on the device
    double *Hd = nullptr, *csd = nullptr, *snd = nullptr, *gd = nullptr, *yd = nullptr, *std_ = nullptr;
    int nblk = 1;
    int zchunk = 4;             // planes per operator workgroup (set by default_shape)
    int colw = 32;              // operator column width (nodes)
    int pred[2] = {0, 0};       // iterations of the last steady / transient solve (launch-ahead count)

    PrecondState<float> p32;
    PrecondState<double> p64;
    bool have_state = false;
    // Per-solve row constants on the device (read by captured launches), and
    // the pinned staging copy they are uploaded from.
    OpK *kdev = nullptr, *kstage = nullptr;
    // One captured graph per FGMRES iteration index and operator (steady /
    // transient); the precision is fixed per handle.
    std::vector<hipGraphExec_t> graphs[2];

    qms_stats st;

    // qms_run_async: the library thread running the steps and its status
    std::thread worker;
    std::atomic<int> running{0};
    // the thread that runs the steps, set by that thread itself before its
    // first call (h->worker is assigned only once the thread has started)
    std::atomic<std::thread::id> owner{};
    int async_rc = QMS_OK;
    std::vector<double> async_v;

    void *raw_alloc(size_t nb) {
        void *p = nullptr;
        check(hipMalloc(&p, nb), "hipMalloc");
        allocs.push_back(p);
        bytes += (int64_t)nb;
        check(hipMemsetAsync(p, 0, nb, s), "hipMemsetAsync");
        return p;
    }
    template <typename T>
    T *dalloc(int64_t count) {
        return (T *)raw_alloc(sizeof(T) * (size_t)(count > 0 ? count : 1));
    }
    ~qms_handle() {
        if (worker.joinable()) worker.join();
        if (s) hipStreamSynchronize(s);
        for (auto &gv : graphs)
            for (hipGraphExec_t e : gv)
                if (e) hipGraphExecDestroy(e);
        for (void *p : allocs) hipFree(p);
        if (hsmall) hipHostFree(hsmall);
        if (kstage) hipHostFree(kstage);
        if (s) hipStreamDestroy(s);
    }
    MgAlloc mg_alloc() {
        return MgAlloc{[this](size_t nb) { return raw_alloc(nb); }, s};
    }
    HostCoef host_coef() const {
        const int64_t n1 = g.n1;
        HostCoef c;
        for (int a = 0; a < 3; ++a) {
            c.im[a] = hcoef.data() + (size_t)a * n1;
            c.sg[a] = hcoef.data() + (size_t)(3 + a) * n1;
        }
        return c;
    }
};

namespace {

// Distinct (1/Mu_x, 1/Mu_y, 1/Mu_z, Sigma_x, Sigma_y, Sigma_z) tuples of the
// design's nodes, keyed by their bit patterns.  Designs built from boxes of a
// few materials have a few hundred; above 65535 the operator keeps streaming
// the six arrays.
struct TupleHash {
    size_t operator()(const std::array<uint64_t, 6> &k) const {
        uint64_t hsh = 0x9e3779b97f4a7c15ull;
        for (uint64_t v : k) hsh = (hsh ^ v) * 0x100000001b3ull + (hsh >> 29);
        return (size_t)hsh;
    }
};

// Also keeps the host copy of the six arrays that the multigrid setup reads.
void build_classes(qms_handle *h) {
    const int64_t n1 = h->g.n1;
    std::vector<double> &a = h->hcoef;
    a.assign((size_t)6 * n1, 0.0);
    for (int q = 0; q < 6; ++q)
        check(hipMemcpyAsync(a.data() + (size_t)q * n1, q < 3 ? h->coef.im[q] : h->coef.sg[q - 3],
                             sizeof(double) * n1, hipMemcpyDeviceToHost, h->s), "d2h");
    check(hipStreamSynchronize(h->s), "sync");
    std::unordered_map<std::array<uint64_t, 6>, uint16_t, TupleHash> map;
    std::vector<uint16_t> cls((size_t)n1);
    std::vector<double> tab;
    for (int64_t p = 0; p < n1; ++p) {
        std::array<uint64_t, 6> key;
        for (int q = 0; q < 6; ++q) memcpy(&key[q], &a[(size_t)q * n1 + p], 8);
        auto it = map.find(key);
        if (it == map.end()) {
            if (map.size() >= 65535) return;           // too many: keep the streamed arrays
            it = map.emplace(key, (uint16_t)map.size()).first;
            for (int q = 0; q < 6; ++q) tab.push_back(a[(size_t)q * n1 + p]);
        }
        cls[(size_t)p] = it->second;
    }
    h->cls = h->dalloc<uint16_t>(n1);
    h->ctab = h->dalloc<double>((int64_t)tab.size());
    check(hipMemcpyAsync(h->cls, cls.data(), sizeof(uint16_t) * n1, hipMemcpyHostToDevice, h->s), "h2d");
    check(hipMemcpyAsync(h->ctab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice, h->s), "h2d");
    check(hipStreamSynchronize(h->s), "sync");
    h->ncls = (int)map.size();
}

// The coefficient view the operator kernels get: with the class table when
// it was built and is enabled.
CoeffsD op_coeffs(const qms_handle *h) {
    CoeffsD c = h->c64;
    if (h->cls && h->use_cls) { c.cls = h->cls; c.ctab = h->ctab; }
    return c;
}

// Default operator shape: 32 x 8 columns, and the deepest z-chunk of
// 1, 2, 4, 8, 16 planes that leaves at least 768 workgroup columns (three
// per CU of 256), so the launch still fills the GPU.
void default_shape(qms_handle *h) {
    h->colw = 32;
    const int64_t cols2d = (int64_t)((h->g.nx + 31) / 32) * ((h->g.ny + 7) / 8);
    int kz = 1;
    for (int k : {2, 4, 8, 16})
        if (cols2d * ((h->g.nz + k - 1) / k) >= 768) kz = k;
    h->zchunk = kz;
}

OpParams make_op(const qms_handle *h, const double v[3], double dt, int tr) {
    OpParams op;
    op.v[0] = v[0]; op.v[1] = v[1]; op.v[2] = v[2];
    op.dt = dt;
    op.tr = tr;
    op.adv = (v[0] != 0.0 || v[1] != 0.0 || v[2] != 0.0) ? 1 : 0;
    op.zchunk = h->zchunk;
    op.colw = h->colw;
    return op;
}

// Multigrid parameters of the A block (block 0) and of the V block (1).
MGParams mgp(const qms_handle *h, int block = 0) {
    MGParams m;
    m.nu = h->cfg.mg_nu;
    m.nu_coarse = h->cfg.mg_nu_coarse > 0 ? h->cfg.mg_nu_coarse : h->cfg.mg_nu;
    m.omega = h->cfg.mg_omega;
    m.alpha = h->cfg.mg_alpha;
    m.tail = h->cfg.mg_fused_tail;
    m.fuse_restrict = h->cfg.mg_fused_restrict;
    m.cheb_lmax = h->cfg.mg_cheb_lmax;
    m.cheb_ratio = h->cfg.mg_cheb_ratio > 1.0 ? h->cfg.mg_cheb_ratio : 4.0;
    if (block == 1) {
        if (h->cfg.mg_nu_v > 0) m.nu = h->cfg.mg_nu_v;
        if (h->cfg.mg_omega_v > 0) m.omega = h->cfg.mg_omega_v;
        if (h->cfg.mg_alpha_v > 0) m.alpha = h->cfg.mg_alpha_v;
        m.cheb_lmax = h->cfg.mg_cheb_lmax_v;
        m.cheb_ratio = h->cfg.mg_cheb_ratio_v > 1.0 ? h->cfg.mg_cheb_ratio_v : 4.0;
    }
    return m;
}

template <typename T> PrecondState<T> &pstate(qms_handle *h);
template <> PrecondState<float> &pstate<float>(qms_handle *h) { return h->p32; }
template <> PrecondState<double> &pstate<double>(qms_handle *h) { return h->p64; }
template <typename T> CoeffsT<T> dev_coef(const qms_handle *h);
template <> CoeffsT<float> dev_coef<float>(const qms_handle *h) {
    return CoeffsT<float>{{h->c32.im[0], h->c32.im[1], h->c32.im[2]}, {h->c32.sg[0], h->c32.sg[1], h->c32.sg[2]},
                          h->use_cls ? h->cls : nullptr, h->use_cls ? h->ctab : nullptr};
}
template <> CoeffsT<double> dev_coef<double>(const qms_handle *h) {
    return CoeffsT<double>{{h->c64.im[0], h->c64.im[1], h->c64.im[2]}, {h->c64.sg[0], h->c64.sg[1], h->c64.sg[2]},
                           h->use_cls ? h->cls : nullptr, h->use_cls ? h->ctab : nullptr};
}

// The Krylov basis of storage type TV (allocated on first use, outside any
// graph capture).  The FP64 basis's w is the handle's work vector.
template <typename TV> KBasis<TV> &kbasis(qms_handle *h);
template <> KBasis<double> &kbasis<double>(qms_handle *h) { return h->kb64; }
template <> KBasis<float> &kbasis<float>(qms_handle *h) { return h->kb32; }

// Byte offset of the V segment of a Krylov vector with A components of type TE
// (256-byte aligned when it does not follow on directly).
template <typename TE>
int64_t seg_offv(int64_t n1) {
    if (std::is_same<TE, double>::value) return 24 * n1;
    return (int64_t)((3 * n1 * (int64_t)sizeof(TE) + 255) / 256 * 256);
}

template <typename TE>
KBasis<TE> &ensure_basis(qms_handle *h) {
    KBasis<TE> &kb = kbasis<TE>(h);
    if (kb.V.empty()) {
        const int m = h->cfg.restart;
        kb.offv = seg_offv<TE>(h->g.n1);
        const size_t nb = (size_t)kb.offv + 8 * (size_t)h->g.n1;
        for (int j = 0; j <= m; ++j) kb.V.push_back((TE *)h->raw_alloc(nb));
        kb.dV = h->dalloc<TE *>(m + 1);
        check(hipMemcpyAsync(kb.dV, kb.V.data(), sizeof(TE *) * (m + 1), hipMemcpyHostToDevice, h->s), "h2d");
        if constexpr (std::is_same<TE, double>::value) kb.w = h->w;
        else kb.w = (TE *)h->raw_alloc(nb);
        check(hipStreamSynchronize(h->s), "sync");
    }
    return kb;
}

// Everything the solve in precision TA needs, allocated and built before any
// kernel of the solve is queued: the multigrid hierarchies (host-built
// Galerkin levels, uploaded), the Krylov basis, the Z basis and the
// preconditioner's work vectors.
template <typename TA>
void ensure_solver(qms_handle *h, int tr) {
    PrecondState<TA> &ps = pstate<TA>(h);
    const HostCoef hc = h->host_coef();
    const MgAlloc al = h->mg_alloc();
    ensure_basis<TA>(h);
    if (ps.Z.empty()) {
        const int m = h->cfg.restart;
        for (int j = 0; j < m; ++j) ps.Z.push_back(h->dalloc<TA>(h->n));
        ps.dZ = h->dalloc<TA *>(m);
        check(hipMemcpyAsync(ps.dZ, ps.Z.data(), sizeof(TA *) * m, hipMemcpyHostToDevice, h->s), "h2d");
        ps.rT = h->dalloc<TA>(3 * h->g.n1);
        ps.rV = h->dalloc<double>(h->g.n1);
        ps.zV = h->dalloc<double>(h->g.n1);
    }
    if (!ps.builtV) {
        mg_setup<double>(h->g, hc, dev_coef<double>(h), 0.0, 1, al, ps.hV);
        ps.builtV = true;
    }
    if (!ps.builtA[tr]) {
        mg_setup<TA>(h->g, hc, dev_coef<TA>(h), tr ? h->cfg.dt : 0.0, 0, al, ps.hA[tr]);
        ps.builtA[tr] = true;
    }
    check(hipStreamSynchronize(h->s), "sync");
    h->st.n_levels_a = ps.hA[tr].nlev;
    h->st.n_levels_v = ps.hV.nlev;
    h->st.mg_classes_a = ps.hA[tr].nlev > 1 ? ps.hA[tr].lev[1].ncls : 0;
    h->st.mg_classes_v = ps.hV.nlev > 1 ? ps.hV.lev[1].ncls : 0;
}

// z = P^-1 v: block lower-triangular [K_A 0; M_VA L_V]^-1 with one V-cycle
// per diagonal block, applied to D v (unscaled residual units).  The V
// block's right-hand side D_V v_V - M_VA z_A is formed in FP64 by the
// operator kernel and its V-cycle runs in FP64.  TA: the A block's multigrid,
// z and the A segment of the Krylov vector v (V segment from byte offv).
template <typename TA>
void precond(qms_handle *h, const OpParams &op, int tr, const double *D, const TA *v, int64_t offv, TA *z) {
    PrecondState<TA> &ps = pstate<TA>(h);
    const int64_t n1 = h->g.n1;
    const MGParams mp = mgp(h);
    // each block's level-0 first sweep from zero is fused into (A) or
    // follows (V) the kernel that forms its right-hand side
    launch_d2t_first(ps.hA[tr], mp, v, D, ps.rT, mg_first_target(ps.hA[tr], mp, z), h->s);
    mg_vcycle<TA>(ps.hA[tr], mp, z, ps.rT, h->s, true);
    if constexpr (std::is_same<TA, double>::value)
        launch_vrhs(h->g, op_coeffs(h), op, z, v, ps.rV, h->s);
    else
        launch_vrhs_seg(h->g, op_coeffs(h), op, z, (const double *)((const char *)v + offv), ps.rV, h->s);
    double *zv = std::is_same<TA, double>::value ? (double *)(void *)(z + 3 * n1) : ps.zV;
    const MGParams mpv = mgp(h, 1);
    launch_first_sweep(ps.hV, mpv, (const double *)ps.rV, mg_first_target(ps.hV, mpv, zv), h->s);
    mg_vcycle<double>(ps.hV, mpv, zv, ps.rV, h->s, true);
    if constexpr (!std::is_same<TA, double>::value) launch_cvt(ps.zV, z + 3 * n1, n1, h->s);
}

// ||w||_2, read back on the host (one synchronisation).
double norm_dev(qms_handle *h, const double *w) {
    launch_norm2(w, h->n, h->partial, h->nblk, h->s);
    launch_reduce_partials(h->partial, 1, h->nblk, h->dsmall, h->s);
    check(hipMemcpyAsync(h->hsmall, h->dsmall, sizeof(double), hipMemcpyDeviceToHost, h->s), "d2h");
    check(hipStreamSynchronize(h->s), "sync");
    return sqrt(h->hsmall[0]);
}

// Classical Gram-Schmidt of kb.w against basis vectors V[0..nv-1]: the dots
// land in dsmall[0..nv-1], w is updated in place and |w|^2 lands in
// dsmall[nv] (one fused pass for the dots, one for the update; FP64 sums).
template <typename TE>
void gram_schmidt(qms_handle *h, KBasis<TE> &kb, int nv) {
    launch_multidot<TE>((const TE *const *)kb.dV, nv, kb.w, h->g.n1, kb.offv, h->partial, h->nblk, h->s);
    launch_reduce_partials(h->partial, nv, h->nblk, h->dsmall, h->s);
    launch_orth_update<TE>(kb.dV, nv, h->dsmall, kb.w, h->g.n1, kb.offv, h->partial, h->nblk, h->s);
    launch_reduce_partials(h->partial, 1, h->nblk, h->dsmall + nv, h->s);
}

void launch_apply_t(qms_handle *h, const OpParams &op, const double *x, double *y, int64_t) {
    launch_apply(h->g, op_coeffs(h), op, x, y, h->s);
}
void launch_apply_t(qms_handle *h, const OpParams &op, const float *x, float *y, int64_t offv) {
    launch_apply_f32(h->g, op_coeffs(h), op, x, y, offv, h->s);
}

// One FGMRES iteration j, queued with no host round trip: z_j = P^-1 v_j,
// w = D^-1 M z_j, Gram-Schmidt, Givens update on the device, v_{j+1}.
template <typename T>
void fgmres_iteration(qms_handle *h, const OpParams &op, int tr, const double *D, int j) {
    PrecondState<T> &ps = pstate<T>(h);
    KBasis<T> &kb = kbasis<T>(h);
    const int m = h->cfg.restart;
    precond<T>(h, op, tr, D, kb.V[j], kb.offv, ps.Z[j]);
    launch_apply_t(h, op, ps.Z[j], kb.w, kb.offv);
    gram_schmidt<T>(h, kb, j + 1);
    launch_givens(j, m, h->dsmall, h->Hd, h->csd, h->snd, h->gd, h->std_, h->s);
    launch_scale_dev<T>(kb.V[j + 1], kb.w, h->std_ + KS_INVH, h->g.n1, kb.offv, h->s);
}

// Iteration j as a captured graph (captured on first use; every argument it
// holds -- buffers, the device row constants, the iteration index -- is the
// same for every later solve of this operator, so replays are the same work).
template <typename T>
void run_iteration(qms_handle *h, const OpParams &op, int tr, const double *D, int j) {
    if (!h->cfg.use_graphs) { fgmres_iteration<T>(h, op, tr, D, j); return; }
    std::vector<hipGraphExec_t> &gv = h->graphs[tr];
    if (gv.empty()) gv.assign(h->cfg.restart, nullptr);
    if (!gv[j]) {
        hipGraph_t g = nullptr;
        check(hipStreamBeginCapture(h->s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        try {
            fgmres_iteration<T>(h, op, tr, D, j);
        } catch (...) {
            hipStreamEndCapture(h->s, &g);
            if (g) hipGraphDestroy(g);
            throw;
        }
        check(hipStreamEndCapture(h->s, &g), "hipStreamEndCapture");
        hipGraphExec_t e = nullptr;
        const hipError_t ie = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        check(ie, "hipGraphInstantiate");
        gv[j] = e;
    }
    check(hipGraphLaunch(gv[j], h->s), "hipGraphLaunch");
}

// Uploads the row constants of this solve to the device copy the solve's
// kernels read (stream-ordered; the staging buffer is rewritten only after the
// previous solve synchronised).
OpParams device_op(qms_handle *h, OpParams op) {
    *h->kstage = make_opk(h->g, op);
    check(hipMemcpyAsync(h->kdev, h->kstage, sizeof(OpK), hipMemcpyHostToDevice, h->s), "h2d");
    op.kdev = h->kdev;
    return op;
}

// Reads the device cycle scalars (one synchronisation).
void read_cycle(qms_handle *h, double out[KS_COUNT]) {
    check(hipMemcpyAsync(h->hsmall, h->std_, sizeof(double) * KS_COUNT, hipMemcpyDeviceToHost, h->s), "d2h");
    check(hipStreamSynchronize(h->s), "sync");
    memcpy(out, h->hsmall, sizeof(double) * KS_COUNT);
}

// Restarted FGMRES on D^-1 M x = D^-1 b with an FP64 residual at every
// restart.  T is the storage precision of the Krylov basis, the Z basis and
// the A block's multigrid:
//   double: plain FP64; a cycle runs until the tolerance is met or the
//           restart length is reached;
//   float:  FP64 iterative refinement around an inner FGMRES whose Krylov
//           vectors hold the A components in FP32 (V in FP64) and whose Z basis
//           and A-block multigrid are FP32: each cycle solves for the
//           correction of the current FP64 residual until that residual is
//           reduced by the factor inner_tol (or the tolerance or the restart
//           length is reached); x accumulates in FP64.
// Iterations are queued ahead in chunks (the first chunk three quarters as
// long as the last cycle), and the host reads the device's convergence record
// only between chunks.
template <typename T>
int fgmres(qms_handle *h, const OpParams &op, int tr) {
    PrecondState<T> &ps = pstate<T>(h);
    KBasis<T> &kb = kbasis<T>(h);
    const int64_t n = h->n;
    const int m = h->cfg.restart;
    const double *D = tr ? h->Dtr : h->Dst;
    hipStream_t s = h->s;
    // ||D^-1 b||: the residual of x = 0
    launch_fill<double>(h->w, n, 0.0, s);
    launch_resid(h->g, op_coeffs(h), op, h->w, h->b, h->r, s);
    const double bnorm = norm_dev(h, h->r);
    const double tol = h->cfg.tol;
    const bool inner = !std::is_same<T, double>::value;
    const double eta = h->cfg.inner_tol > 0 ? h->cfg.inner_tol : 1e-6;
    const double floor_tol = std::max(tol, h->cfg.floor_tol > 0 ? h->cfg.floor_tol : 100.0 * tol);
    int used = 0, queued = 0, restarts = 0, stall = 0;
    double rel = 0.0, rel0 = -1.0;
    bool cycle_stalled = false;
    double cyc[KS_COUNT];
    while (true) {
        launch_resid(h->g, op_coeffs(h), op, h->x, h->b, h->r, s);
        const double beta = norm_dev(h, h->r);
        rel = bnorm > 0 ? beta / bnorm : beta;
        if (rel0 < 0) rel0 = rel;
        if (!(rel == rel)) throw HipError("residual is NaN", QMS_ERR_NOCONV);
        if (rel <= tol || beta == 0.0) break;
        // Attainable accuracy: the FP64 residual of a step has a floor set by
        // rounding in rows whose scales span 1e-3..1e15, and on the config-3
        // ramp it rises from about 1e-13 to 1e-11 as the fields grow; near it
        // FGMRES stagnates, whatever the restart length.  A cycle that
        // stopped improving (its own estimate, checked on the device over the
        // last two iterations, or one that ended above half the residual it
        // started from) is accepted when the FP64 residual is within
        // floor_tol (and counted); two such cycles in a row above it fail at
        // once instead of running to the iteration cap.
        if (cycle_stalled && rel <= floor_tol) { h->st.stalled_solves += 1; break; }
        stall = cycle_stalled ? stall + 1 : 0;
        if (stall >= 2) {
            h->st.last_iters = used; h->st.last_restarts = restarts; h->st.last_relres = rel;
            h->st.last_relres0 = rel0;
            char buf[160];
            snprintf(buf, sizeof buf, "FGMRES stalled after %d iterations at relative residual %.3e (tol %.1e)",
                     used, rel, tol);
            throw HipError(buf, QMS_ERR_NOCONV);
        }
        if (used >= h->cfg.max_iters) {
            h->st.last_iters = used; h->st.last_restarts = restarts; h->st.last_relres = rel;
            h->st.last_relres0 = rel0;
            char buf[160];
            snprintf(buf, sizeof buf, "FGMRES stopped after %d iterations at relative residual %.3e (tol %.1e)",
                     used, rel, tol);
            throw HipError(buf, QMS_ERR_NOCONV);
        }
        ++restarts;
        launch_scale<T>(kb.V[0], h->r, 1.0 / beta, h->g.n1, kb.offv, s);
        // the cycle stops at an estimated residual of half its target
        const double target = inner ? std::max(tol, 2.0 * eta * rel) : tol;
        launch_cycle_start(m, beta, bnorm, target, floor_tol, h->gd, h->std_, s);
        int j = 0, k = 0;
        // first chunk: three quarters of the last cycle's count (the count
        // varies by a few iterations between steps), then two at a time
        int chunk = h->pred[tr] > 0 ? std::max(2, std::min(3 * h->pred[tr] / 4, m)) : 4;
        while (j < m) {
            const int nq = std::min(chunk, m - j);
            for (int q = 0; q < nq; ++q) run_iteration<T>(h, op, tr, D, j + q);
            j += nq;
            queued += nq;
            read_cycle(h, cyc);
            if (!(cyc[KS_EST] == cyc[KS_EST])) throw HipError("FGMRES residual estimate is NaN", QMS_ERR_NOCONV);
            if (cyc[KS_K] > 0) { k = (int)cyc[KS_K]; break; }
            chunk = 2;
        }
        if (k == 0) k = m;
        cycle_stalled = cyc[KS_STALL] > 0 || cyc[KS_EST] > 0.5 * rel;
        used += k;
        launch_backsub(k, m, h->Hd, h->gd, h->yd, s);
        launch_axpy_list(h->x, (const T *const *)ps.dZ, k, h->yd, n, s);
        h->pred[tr] = k;
    }
    h->st.last_iters = used;
    h->st.last_restarts = restarts;
    h->st.last_relres = rel;
    h->st.last_relres0 = rel0;
    h->st.total_iters += used;
    h->st.solves += 1;
    h->st.queued_iters += queued;
    return QMS_OK;
}

int solve(qms_handle *h, const OpParams &op0, int tr) {
    const OpParams op = device_op(h, op0);
    if (h->cfg.precision == QMS_MIXED) return fgmres<float>(h, op, tr);
    return fgmres<double>(h, op, tr);
}

void prepare_solve(qms_handle *h, int tr) {
    if (h->cfg.precision == QMS_MIXED) ensure_solver<float>(h, tr);
    else ensure_solver<double>(h, tr);
}

// True (and the handle's message set) when another thread than the handle's
// asynchronous run calls while that run is going.
bool busy(qms_handle *h) {
    if (h->running.load() && std::this_thread::get_id() != h->owner.load()) {
        h->err = "handle busy: an asynchronous run is going (qms_wait first)";
        return true;
    }
    return false;
}

int fail(qms_handle *h, const std::exception &e, int code) {
    if (h) h->err = e.what();
    g_last_error = e.what();
    return code;
}

}  // namespace

#define QMS_TRY(h) try {
#define QMS_CATCH(h)                                                   \
    }                                                                  \
    catch (const HipError &e) { return fail(h, e, e.code); }           \
    catch (const std::bad_alloc &e) { return fail(h, e, QMS_ERR_NOMEM); } \
    catch (const std::exception &e) { return fail(h, e, QMS_ERR_ARG); }

extern "C" {

int qms_config_default(qms_config *cfg) {
    if (!cfg) return QMS_ERR_ARG;
    memset(cfg, 0, sizeof(*cfg));
    cfg->precision = QMS_FP64;
    cfg->tol = 1e-12;
    cfg->restart = 30;
    cfg->max_iters = 3000;
    cfg->mg_nu = 4;
    cfg->mg_omega = 0.7;
    cfg->mg_alpha = 1.5;
    cfg->use_graphs = 1;
    cfg->inner_tol = 1e-6;
    cfg->mg_nu_coarse = 3;
    cfg->floor_tol = 0.0;
    cfg->mg_fused_tail = 1;
    cfg->mg_omega_v = 0.85;
    cfg->mg_nu_v = 6;
    cfg->mg_alpha_v = 1.9;
    cfg->mg_cheb_lmax = 2.6;
    cfg->mg_cheb_ratio = 4.0;
    cfg->mg_cheb_lmax_v = 2.0;
    cfg->mg_cheb_ratio_v = 8.0;
    cfg->mg_fused_restrict = 1;
    return QMS_OK;
}

int qms_create(const qms_config *cfg, const uint16_t *cell_slot, const double *slot_props, int nslots,
               qms_handle **out) {
    if (!cfg || !cell_slot || !slot_props || !out || nslots < 1) {
        g_last_error = "qms_create: null argument or no material slots";
        return QMS_ERR_ARG;
    }
    *out = nullptr;
    for (int a = 0; a < 3; ++a) {
        if (cfg->cells[a] < 5 || cfg->cells[a] > 4096) {
            g_last_error = "qms_create: cells per axis must be in [5, 4096]";
            return QMS_ERR_ARG;
        }
        if (!(cfg->h[a] > 0)) { g_last_error = "qms_create: grid steps must be positive"; return QMS_ERR_ARG; }
    }
    if (cfg->restart < 1 || cfg->restart > 64 || cfg->mg_nu < 1 || cfg->mg_nu > 64 || cfg->mg_nu_coarse > 64 ||
        cfg->mg_nu_v > 64 || !(cfg->mg_omega_v <= 1.0) || !(cfg->tol > 0) ||
        !(cfg->mg_cheb_lmax < 100.0) || !(cfg->mg_cheb_lmax_v < 100.0) || !(cfg->mg_cheb_ratio < 1e6) ||
        !(cfg->mg_cheb_ratio_v < 1e6) ||
        cfg->max_iters < 1 ||
        (cfg->precision != QMS_FP64 && cfg->precision != QMS_MIXED)) {
        g_last_error = "qms_create: invalid solver parameters";
        return QMS_ERR_ARG;
    }
    std::unique_ptr<qms_handle> h(new qms_handle());
    QMS_TRY(h.get())
    h->cfg = *cfg;
    memset(&h->st, 0, sizeof(h->st));
    Grid &g = h->g;
    g.nx = cfg->cells[0] + 1; g.ny = cfg->cells[1] + 1; g.nz = cfg->cells[2] + 1;
    g.n1 = (int64_t)g.nx * g.ny * g.nz;
    g.hx = cfg->h[0]; g.hy = cfg->h[1]; g.hz = cfg->h[2];
    for (int a = 0; a < 3; ++a) h->bc.B[a] = cfg->bc[a];
    for (int a = 0; a < 9; ++a) h->bc.G[a] = cfg->bc[3 + a];
    check(hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking), "hipStreamCreate");
    const int64_t nc = (int64_t)cfg->cells[0] * cfg->cells[1] * cfg->cells[2];
    for (int64_t p = 0; p < nc; ++p)
        if (cell_slot[p] >= nslots) throw std::runtime_error("qms_create: cell slot outside the material table");
    h->slot = h->dalloc<uint16_t>(nc);
    h->mat = h->dalloc<double>(6 * (int64_t)nslots);
    check(hipMemcpyAsync(h->slot, cell_slot, sizeof(uint16_t) * nc, hipMemcpyHostToDevice, h->s), "h2d");
    check(hipMemcpyAsync(h->mat, slot_props, sizeof(double) * 6 * nslots, hipMemcpyHostToDevice, h->s), "h2d");
    for (int a = 0; a < 3; ++a) {
        h->coef.im[a] = h->dalloc<double>(g.n1);
        h->coef.sg[a] = h->dalloc<double>(g.n1);
        h->coef.im32[a] = h->dalloc<float>(g.n1);
        h->coef.sg32[a] = h->dalloc<float>(g.n1);
        h->c64.im[a] = h->coef.im[a]; h->c64.sg[a] = h->coef.sg[a];
        h->c32.im[a] = h->coef.im32[a]; h->c32.sg[a] = h->coef.sg32[a];
    }
    h->c64.cls = nullptr;
    h->c64.ctab = nullptr;
    launch_coeffs(g, h->slot, h->mat, h->coef, h->s);
    build_classes(h.get());
    h->n = 4 * g.n1;
    default_shape(h.get());
    h->x = h->dalloc<double>(h->n);
    h->b = h->dalloc<double>(h->n);
    h->r = h->dalloc<double>(h->n);
    h->w = h->dalloc<double>(h->n);
    h->tmp = h->dalloc<double>(h->n);
    h->Dst = h->dalloc<double>(h->n);
    h->Dtr = h->dalloc<double>(h->n);
    h->xh1 = h->dalloc<double>(h->n);
    h->xh2 = h->dalloc<double>(h->n);
    const double v0[3] = {0, 0, 0};
    launch_diag(g, h->c64, make_op(h.get(), v0, 0.0, 0), h->Dst, h->s);
    launch_diag(g, h->c64, make_op(h.get(), v0, cfg->dt, cfg->dt > 0 ? 1 : 0), h->Dtr, h->s);
    const int m = cfg->restart;
    h->nblk = reduce_blocks(h->n);
    h->partial = h->dalloc<double>((int64_t)(m + 2) * h->nblk);
    h->dsmall = h->dalloc<double>(m + 8);
    h->Hd = h->dalloc<double>((int64_t)(m + 1) * m);
    h->csd = h->dalloc<double>(m);
    h->snd = h->dalloc<double>(m);
    h->gd = h->dalloc<double>(m + 1);
    h->yd = h->dalloc<double>(m);
    h->std_ = h->dalloc<double>(KS_COUNT);
    h->kdev = h->dalloc<OpK>(1);
    check(hipHostMalloc((void **)&h->kstage, sizeof(OpK), hipHostMallocDefault), "hipHostMalloc");
    check(hipHostMalloc((void **)&h->hsmall, sizeof(double) * (m + 8 > KS_COUNT ? m + 8 : KS_COUNT),
                        hipHostMallocDefault), "hipHostMalloc");
    check(hipStreamSynchronize(h->s), "sync");
    h->st.device_bytes = h->bytes;
    *out = h.release();
    return QMS_OK;
    QMS_CATCH(h.get())
}

int qms_steady(qms_handle *h, const double v[3]) {
    if (!h || !v) return QMS_ERR_ARG;
    if (busy(h)) return QMS_ERR_STATE;
    QMS_TRY(h)
    prepare_solve(h, 0);
    const OpParams op = make_op(h, v, 0.0, 0);
    launch_rhs(h->g, op_coeffs(h), h->bc, op, h->x, h->b, h->s);
    launch_fill<double>(h->x, h->n, 0.0, h->s);
    launch_setbc(h->g, h->bc, op, h->x, h->s);
    h->nhist = 0;
    solve(h, op, 0);
    h->have_state = true;
    return QMS_OK;
    QMS_CATCH(h)
}

int qms_step(qms_handle *h, const double v[3]) {
    if (!h || !v) return QMS_ERR_ARG;
    if (busy(h)) return QMS_ERR_STATE;
    if (!h->have_state) { h->err = "qms_step: no state (call qms_steady or qms_set_fields first)"; return QMS_ERR_STATE; }
    if (!(h->cfg.dt > 0)) { h->err = "qms_step: dt must be positive"; return QMS_ERR_ARG; }
    QMS_TRY(h)
    prepare_solve(h, 1);
    const OpParams op = make_op(h, v, h->cfg.dt, 1);
    launch_rhs(h->g, op_coeffs(h), h->bc, op, h->x, h->b, h->s);
    // Initial guess: the solution extrapolated in time from the last steps
    // (exact for a solution that moves linearly or quadratically in time,
    // as under the reference's constant-acceleration trajectory).
    launch_extrapolate(h->x, h->xh1, h->xh2, h->nhist, h->n, h->s);
    h->nhist = std::min(h->nhist + 1, 2);
    launch_setbc(h->g, h->bc, op, h->x, h->s);
    solve(h, op, 1);
    return QMS_OK;
    QMS_CATCH(h)
}

int qms_run(qms_handle *h, int nsteps, const double *v) {
    if (!h || (nsteps > 0 && !v) || nsteps < 0) return QMS_ERR_ARG;
    if (busy(h)) return QMS_ERR_STATE;
    for (int sidx = 0; sidx < nsteps; ++sidx) {
        const int rc = qms_step(h, v + 3 * sidx);
        if (rc != QMS_OK) return rc;
    }
    return QMS_OK;
}

int qms_run_async(qms_handle *h, int nsteps, const double *v) {
    if (!h || (nsteps > 0 && !v) || nsteps < 0) return QMS_ERR_ARG;
    if (h->running.load()) { h->err = "qms_run_async: a run is already going on this handle"; return QMS_ERR_STATE; }
    if (h->worker.joinable()) h->worker.join();
    h->async_v.assign(v, v + 3 * (size_t)nsteps);
    h->async_rc = QMS_OK;
    h->running.store(1);
    try {
        h->worker = std::thread([h, nsteps]() {
            h->owner.store(std::this_thread::get_id());
            h->async_rc = qms_run(h, nsteps, h->async_v.data());
            h->running.store(0);
        });
    } catch (const std::exception &e) {
        h->running.store(0);
        return fail(h, e, QMS_ERR_ARG);
    }
    return QMS_OK;
}

int qms_poll(qms_handle *h) {
    if (!h) return QMS_ERR_ARG;
    if (h->running.load()) return QMS_BUSY;
    return h->async_rc;
}

int qms_wait(qms_handle *h) {
    if (!h) return QMS_ERR_ARG;
    if (h->worker.joinable()) h->worker.join();
    return h->async_rc;
}

int64_t qms_estimate_bytes(const qms_config *cfg) {
    if (!cfg) return 0;
    for (int a = 0; a < 3; ++a)
        if (cfg->cells[a] < 5 || cfg->cells[a] > 4096) return 0;
    if (cfg->restart < 1 || cfg->restart > 64) return 0;
    const int64_t nx = cfg->cells[0] + 1, ny = cfg->cells[1] + 1, nz = cfg->cells[2] + 1;
    const int64_t n1 = nx * ny * nz, n = 4 * n1, nc = (int64_t)cfg->cells[0] * cfg->cells[1] * cfg->cells[2];
    const int64_t m = cfg->restart;
    const int64_t ta = cfg->precision == QMS_MIXED ? 4 : 8;
    int64_t b = 0, allocs = 0;
    auto add = [&](int64_t bytes) { b += bytes; ++allocs; };
    // qms_create: cell slots, material table (at most 65536 slots), coefficients in FP64 and FP32,
    // class indices and table, state and work vectors, row scalings, step history
    add(2 * nc); add(48 * 65536);
    for (int a = 0; a < 6; ++a) { add(8 * n1); add(4 * n1); }
    add(2 * n1); add(48 * 65535);
    for (int v = 0; v < 9; ++v) add(8 * n);
    add(8 * (m + 2) * std::max<int64_t>(1, std::min<int64_t>(1024, (n + 2047) / 2048)));
    add(8 * (m + 8)); add(8 * (m + 1) * m); add(8 * m); add(8 * m); add(8 * (m + 1)); add(8 * m);
    add(8 * KS_COUNT); add((int64_t)sizeof(OpK));
    // first solve: Krylov basis and its w (precision ta), Z basis and preconditioner vectors
    const int64_t kvec = (3 * n1 * ta + 255) / 256 * 256 + 8 * n1;    // A segment in precision ta, V in FP64
    for (int j = 0; j <= m; ++j) add(kvec);
    add(8 * (m + 1));
    if (ta != 8) add(kvec);
    for (int j = 0; j < m; ++j) add(ta * n);
    add(8 * m); add(ta * 3 * n1); add(8 * n1); add(8 * n1);
    // hierarchies: A block steady and transient (precision ta), V block (FP64)
    auto hier = [&](int nc_, int64_t s) {
        int64_t x = nx, y = ny, z = nz;
        add(s * nc_ * x * y * z); add(s * nc_ * x * y * z);     // level 0: t, r
        add(s * nc_ * x * y * z);                               // level 0: diagonal
        int nlev = 1;
        int64_t cn = 0;
        while (nlev < MG_MAX_LEVELS) {
            if (std::max(x, std::max(y, z)) <= 4 && nlev >= 2) break;
            x = (x + 1) / 2; y = (y + 1) / 2; z = (z + 1) / 2;
            const int64_t l1 = x * y * z;
            // explicit stencil arrays (an upper bound of the class-compressed form)
            add(s * nc_ * 6 * l1); add(s * nc_ * l1); add(s * nc_ * l1);
            if (nc_ == 3) add(s * nc_ * 8 * l1);
            for (int v = 0; v < 4; ++v) add(s * nc_ * l1);
            cn = nc_ * l1;
            ++nlev;
        }
        add(4 * cn); add(8 * cn * cn);                              // coarsest dense inverse (upper bound)
    };
    hier(3, ta);
    hier(3, ta);
    hier(1, 8);
    add(2 * n1); add(8 * 8 * 8192);                             // V block: level-0 row classes and table
    return b + allocs * 256;       // allocation granularity
}

int qms_get_fields(qms_handle *h, double *out) {
    if (!h || !out) return QMS_ERR_ARG;
    if (busy(h)) return QMS_ERR_STATE;
    QMS_TRY(h)
    check(hipMemcpyAsync(out, h->x, sizeof(double) * h->n, hipMemcpyDeviceToHost, h->s), "d2h");
    check(hipStreamSynchronize(h->s), "sync");
    return QMS_OK;
    QMS_CATCH(h)
}

int qms_set_fields(qms_handle *h, const double *in) {
    if (!h || !in) return QMS_ERR_ARG;
    if (busy(h)) return QMS_ERR_STATE;
    QMS_TRY(h)
    check(hipMemcpyAsync(h->x, in, sizeof(double) * h->n, hipMemcpyHostToDevice, h->s), "h2d");
    check(hipStreamSynchronize(h->s), "sync");
    h->have_state = true;
    h->nhist = 0;
    return QMS_OK;
    QMS_CATCH(h)
}

int qms_get_h(qms_handle *h, int unit_mu, double *out) {
    if (!h || !out) return QMS_ERR_ARG;
    if (busy(h)) return QMS_ERR_STATE;
    QMS_TRY(h)
    const int64_t nc = (int64_t)h->cfg.cells[0] * h->cfg.cells[1] * h->cfg.cells[2];
    launch_hfield(h->g, op_coeffs(h), h->x, h->tmp, unit_mu, h->s);
    check(hipMemcpyAsync(out, h->tmp, sizeof(double) * 3 * nc, hipMemcpyDeviceToHost, h->s), "d2h");
    check(hipStreamSynchronize(h->s), "sync");
    return QMS_OK;
    QMS_CATCH(h)
}

int qms_apply(qms_handle *h, const double v[3], double dt, const double *x, double *y) {
    if (!h || !v || !x || !y) return QMS_ERR_ARG;
    if (busy(h)) return QMS_ERR_STATE;
    QMS_TRY(h)
    const int tr = dt > 0 ? 1 : 0;
    const OpParams op = make_op(h, v, dt, tr);
    check(hipMemcpyAsync(h->r, x, sizeof(double) * h->n, hipMemcpyHostToDevice, h->s), "h2d");
    launch_apply(h->g, op_coeffs(h), op, h->r, h->w, h->s);
    launch_diag(h->g, h->c64, op, h->tmp, h->s);
    check(hipMemcpyAsync(y, h->w, sizeof(double) * h->n, hipMemcpyDeviceToHost, h->s), "d2h");
    std::vector<double> D(h->n);
    check(hipMemcpyAsync(D.data(), h->tmp, sizeof(double) * h->n, hipMemcpyDeviceToHost, h->s), "d2h");
    check(hipStreamSynchronize(h->s), "sync");
    for (int64_t q = 0; q < h->n; ++q) y[q] *= D[q];
    return QMS_OK;
    QMS_CATCH(h)
}

int qms_residual(qms_handle *h, const double v[3], double dt, const double *x, const double *b, double *r) {
    if (!h || !v || !x || !b || !r) return QMS_ERR_ARG;
    if (busy(h)) return QMS_ERR_STATE;
    QMS_TRY(h)
    const int tr = dt > 0 ? 1 : 0;
    const OpParams op = make_op(h, v, dt, tr);
    check(hipMemcpyAsync(h->r, x, sizeof(double) * h->n, hipMemcpyHostToDevice, h->s), "h2d");
    check(hipMemcpyAsync(h->tmp, b, sizeof(double) * h->n, hipMemcpyHostToDevice, h->s), "h2d");
    launch_resid(h->g, op_coeffs(h), op, h->r, h->tmp, h->w, h->s);
    check(hipMemcpyAsync(r, h->w, sizeof(double) * h->n, hipMemcpyDeviceToHost, h->s), "d2h");
    check(hipStreamSynchronize(h->s), "sync");
    return QMS_OK;
    QMS_CATCH(h)
}

int qms_rhs(qms_handle *h, const double v[3], double dt, const double *aprev, double *b) {
    if (!h || !v || !aprev || !b) return QMS_ERR_ARG;
    if (busy(h)) return QMS_ERR_STATE;
    QMS_TRY(h)
    const OpParams op = make_op(h, v, dt, dt > 0 ? 1 : 0);
    check(hipMemcpyAsync(h->r, aprev, sizeof(double) * h->n, hipMemcpyHostToDevice, h->s), "h2d");
    launch_rhs(h->g, op_coeffs(h), h->bc, op, h->r, h->w, h->s);
    check(hipMemcpyAsync(b, h->w, sizeof(double) * h->n, hipMemcpyDeviceToHost, h->s), "d2h");
    check(hipStreamSynchronize(h->s), "sync");
    return QMS_OK;
    QMS_CATCH(h)
}

int qms_get_stats(const qms_handle *h, qms_stats *st) {
    if (!h || !st) return QMS_ERR_ARG;
    *st = h->st;
    st->device_bytes = h->bytes;
    st->coef_classes = h->ncls;
    auto fused = [&](const auto &ps) {
        int nb = ps.builtV && ps.hV.fused_restrict ? 1 : 0;
        if ((ps.builtA[0] && ps.hA[0].fused_restrict) || (ps.builtA[1] && ps.hA[1].fused_restrict)) ++nb;
        return nb;
    };
    st->fused_restrict_blocks = h->cfg.precision == QMS_MIXED ? fused(h->p32) : fused(h->p64);
    return QMS_OK;
}

int qms_set_zchunk(qms_handle *h, int planes) {
    if (!h) return QMS_ERR_ARG;
    if (planes < 1 || planes > 64) {
        h->err = "qms_set_zchunk: planes per workgroup must be in [1, 64]";
        return QMS_ERR_ARG;
    }
    h->zchunk = planes;
    return QMS_OK;
}

int qms_use_coef_classes(qms_handle *h, int on) {
    if (!h) return QMS_ERR_ARG;
    h->use_cls = on != 0;
    return QMS_OK;
}

int qms_set_colwidth(qms_handle *h, int nodes) {
    if (!h) return QMS_ERR_ARG;
    if (nodes != 32 && nodes != 64) {
        h->err = "qms_set_colwidth: column width must be 32 or 64 nodes";
        return QMS_ERR_ARG;
    }
    h->colw = nodes;
    return QMS_OK;
}

int qms_time_kernel(qms_handle *h, int which, const double v[3], int reps, double *ms_per_launch,
                    double *bytes_per_launch) {
    if (!h || reps < 1 || !ms_per_launch || !bytes_per_launch) return QMS_ERR_ARG;
    if (busy(h)) return QMS_ERR_STATE;
    const bool cold = (which & QMS_TIME_COLD) != 0;
    which &= ~QMS_TIME_COLD;
    if (which < 0 || which > 11) return QMS_ERR_ARG;
    QMS_TRY(h)
    if (which == 11) {
        // The copy kernel on two arrays of 2^28 doubles (2 GiB each), far
        // beyond every cache: the sustained HBM rate of a plain stream, with
        // launch and ramp a negligible share of each launch.
        const int64_t big = (int64_t)1 << 28;
        struct Pair {
            hipStream_t s;
            double *p[2] = {nullptr, nullptr};
            ~Pair() {
                hipStreamSynchronize(s);          // no launch may still use them (also on an error exit)
                for (double *q : p) hipFree(q);
            }
        } pr{h->s};
        for (double *&q : pr.p) check(hipMalloc(&q, sizeof(double) * big), "hipMalloc");
        check(hipMemsetAsync(pr.p[1], 0, sizeof(double) * big, h->s), "memset");
        auto launch = [&]() { launch_copy(pr.p[0], pr.p[1], big, h->s); };
        launch();
        *ms_per_launch = time_on(h->s, reps, launch);
        *bytes_per_launch = 16.0 * (double)big;
        return QMS_OK;
    }
    const int tr = h->cfg.dt > 0 ? 1 : 0;
    const double v0[3] = {0, 0, 0};
    const OpParams op = make_op(h, v ? v : v0, h->cfg.dt, tr);
    const int64_t n1 = h->g.n1, n = h->n;
    const int64_t nc = (int64_t)h->cfg.cells[0] * h->cfg.cells[1] * h->cfg.cells[2];
    const bool mixed = h->cfg.precision == QMS_MIXED;
    const double s = mixed ? 4.0 : 8.0;
    if (which == 1 || which == 2 || which >= 8) prepare_solve(h, tr);
    if (which == 7) ensure_basis<double>(h);
    const OpParams opd = which == 10 ? device_op(h, op) : op;
    // The arrays a launch reads (x, b) and writes (w): the handle's own, or
    // one of the copies the cold timing rotates through.
    struct Arrays {
        const double *x, *b;
        double *w;
    };
    Arrays cur{h->x, h->b, which == 5 ? h->tmp : h->w};
    auto launch = [&]() {
        if (which == 0) {
            launch_apply(h->g, op_coeffs(h), op, cur.x, cur.w, h->s);
        } else if (which == 3) {
            launch_copy(cur.w, cur.x, n, h->s);
        } else if (which == 4) {
            launch_rhs(h->g, op_coeffs(h), h->bc, op, cur.x, cur.w, h->s);
        } else if (which == 5) {
            launch_hfield(h->g, op_coeffs(h), cur.x, cur.w, 0, h->s);
        } else if (which == 6) {
            launch_resid(h->g, op_coeffs(h), op, cur.x, cur.b, cur.w, h->s);
        } else if (which == 7) {
            gram_schmidt<double>(h, h->kb64, h->cfg.restart + 1);
        } else if (which >= 8) {
            // a whole A-block or V-block V-cycle, or one FGMRES iteration
            // (iteration index min(10, restart - 1): preconditioner, operator,
            // Gram-Schmidt against up to 11 vectors, Givens, next basis vector)
            const MGParams mp = mgp(h);
            const double *D = tr ? h->Dtr : h->Dst;
            const int j = std::min(10, h->cfg.restart - 1);
            auto part = [&](auto &ps, auto tag) {
                using T = decltype(tag);
                if (which == 8) mg_vcycle<T>(ps.hA[tr], mp, ps.Z[0], ps.rT, h->s);
                else if (which == 9) mg_vcycle<double>(ps.hV, mgp(h, 1), ps.zV, ps.rV, h->s);
                else fgmres_iteration<T>(h, opd, tr, D, j);
            };
            if (mixed) part(h->p32, 0.0f);
            else part(h->p64, 0.0);
        } else {
            const MGParams mp = mgp(h);
            auto sweep = [&](auto &ps) {
                if (which == 1) mg_fine_sweep(ps.hA[tr], mp, ps.Z[0], ps.rT, ps.hA[tr].lev[0].t, h->s);
                else mg_fine_sweep(ps.hV, mgp(h, 1), ps.zV, ps.rV, ps.hV.lev[0].t, h->s);
            };
            if (mixed) sweep(h->p32);
            else sweep(h->p64);
        }
    };
    if (cold && (which == 0 || (which >= 3 && which <= 6))) {
        // Cold caches, as inside a solve, where the Krylov basis streams
        // through the caches between two launches of a kernel: the launches
        // rotate through sets of copies of the kernel's arrays, enough sets
        // that the others' traffic since a set's last use (at least 640 MB)
        // exceeds the L2s and the 256 MB Infinity Cache together.  Launches
        // run back to back and are timed as a whole, like the warm figure.
        // (The Gram-Schmidt step streams 4 GB and is cold as it is; the
        // smoother sweeps are timed as they are.)  Arrays every set shares
        // -- the coefficient classes and table of the operator (4 MB at
        // 128^3), the three sigma arrays the RHS reads (50 MB) -- are read
        // by every launch and stay cached: those parts are warm.
        const int64_t per = (which == 6 ? 24 : 16) * n;          // x (and b) read, the output written, bytes
        const int nsets = (int)std::min<int64_t>(64, 1 + ((int64_t)640 << 20) / per + 1);
        struct Sets {
            hipStream_t s;
            std::vector<double *> p;
            ~Sets() {
                hipStreamSynchronize(s);          // no launch may still use a copy (also on an error exit)
                for (double *q : p) hipFree(q);
            }
        } sets{h->s, {}};
        std::vector<Arrays> arr{cur};
        for (int k = 1; k < nsets; ++k) {
            double *q[3] = {nullptr, nullptr, nullptr};
            for (double *&qq : q) {
                check(hipMalloc(&qq, sizeof(double) * n), "hipMalloc");
                sets.p.push_back(qq);
            }
            check(hipMemcpyAsync(q[0], h->x, sizeof(double) * n, hipMemcpyDeviceToDevice, h->s), "d2d");
            check(hipMemcpyAsync(q[1], h->b, sizeof(double) * n, hipMemcpyDeviceToDevice, h->s), "d2d");
            arr.push_back(Arrays{q[0], q[1], q[2]});
        }
        int next = 0;
        auto rotate = [&]() {
            cur = arr[next];
            next = (next + 1) % nsets;
            launch();
        };
        for (int k = 0; k < nsets; ++k) rotate();               // every set once, untimed
        *ms_per_launch = time_on(h->s, reps, rotate);
    } else {
        *ms_per_launch = time_on(h->s, reps, launch);
    }
    // Algorithmic bytes: each input and output array streamed once.
    const double N = (double)n, N1 = (double)n1, NC = (double)nc;
    // Operator: the workload's algorithmic bytes, x and y (and b) plus the six
    // FP64 coefficient arrays of the reference's operator, whichever way the
    // kernel reads its coefficients.
    switch (which) {
    case 0: *bytes_per_launch = 8.0 * (2.0 * N + 6.0 * N1); break;           // x, y, 6 coefficient arrays
    case 6: *bytes_per_launch = 8.0 * (3.0 * N + 6.0 * N1); break;           // x, b, r, 6 coefficient arrays
    case 3: *bytes_per_launch = 16.0 * N; break;                             // read + write of the state
    case 4: *bytes_per_launch = 8.0 * (3.0 * N1 + 3.0 * N1 + 4.0 * N1); break; // A^n, sigma, b
    case 5: *bytes_per_launch = 8.0 * (6.0 * N1 + 3.0 * NC); break;          // A, 1/Mu, H
    case 7: {                                                                 // dots: m+1 basis + w; update: m+1 basis + w read + w written
        const double m1 = (double)(h->cfg.restart + 1);
        *bytes_per_launch = 8.0 * N * ((m1 + 1.0) + (m1 + 2.0));
        break;
    }
    case 1: *bytes_per_launch = s * (3.0 * 3.0 * N1 + 6.0 * N1); break;
    case 8: case 9: case 10: *bytes_per_launch = 0.0; break;             // composite: no single stream
    default: *bytes_per_launch = s * (3.0 * N1 + 3.0 * N1); break;
    }
    return QMS_OK;
    QMS_CATCH(h)
}

int qms_gram_schmidt(qms_handle *h, int nv, const double *basis, const double *w, double *hcoef, double *wout) {
    if (!h || !basis || !w || !hcoef || !wout) return QMS_ERR_ARG;
    if (busy(h)) return QMS_ERR_STATE;
    if (nv < 1 || nv > h->cfg.restart + 1) {
        h->err = "qms_gram_schmidt: nv must be in [1, restart + 1]";
        return QMS_ERR_ARG;
    }
    QMS_TRY(h)
    const int64_t n = h->n;
    KBasis<double> &kb = ensure_basis<double>(h);
    for (int i = 0; i < nv; ++i)
        check(hipMemcpyAsync(kb.V[i], basis + (size_t)i * n, sizeof(double) * n, hipMemcpyHostToDevice, h->s), "h2d");
    check(hipMemcpyAsync(kb.w, w, sizeof(double) * n, hipMemcpyHostToDevice, h->s), "h2d");
    gram_schmidt<double>(h, kb, nv);
    check(hipMemcpyAsync(h->hsmall, h->dsmall, sizeof(double) * (nv + 1), hipMemcpyDeviceToHost, h->s), "d2h");
    check(hipMemcpyAsync(wout, h->w, sizeof(double) * n, hipMemcpyDeviceToHost, h->s), "d2h");
    check(hipStreamSynchronize(h->s), "sync");
    memcpy(hcoef, h->hsmall, sizeof(double) * (nv + 1));
    return QMS_OK;
    QMS_CATCH(h)
}

int qms_get_coeffs(qms_handle *h, double *out) {
    if (!h || !out) return QMS_ERR_ARG;
    if (busy(h)) return QMS_ERR_STATE;
    const int64_t n1 = h->g.n1;
    for (int q = 0; q < 6; ++q)
        for (int64_t p = 0; p < n1; ++p) {
            const double c = h->hcoef[(size_t)q * n1 + p];
            out[q * n1 + p] = q < 3 ? 1.0 / c : c;
        }
    return QMS_OK;
}

int qms_get_work(qms_handle *h, double *out) {
    if (!h || !out) return QMS_ERR_ARG;
    if (busy(h)) return QMS_ERR_STATE;
    QMS_TRY(h)
    check(hipMemcpyAsync(out, h->w, sizeof(double) * h->n, hipMemcpyDeviceToHost, h->s), "d2h");
    check(hipStreamSynchronize(h->s), "sync");
    return QMS_OK;
    QMS_CATCH(h)
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batches: one operator launch over many designs.
// ---------------------------------------------------------------------------
struct qms_batch {
    std::vector<qms_handle *> hs;
    hipStream_t s = nullptr;
    BatchItem *items_dev = nullptr;
    std::vector<BatchItem> items;
    int blocks = 0;            // persistent grid size, 0 = one workgroup per natural piece
    int tx = 0, cls = 0;       // column width / class-table use shared by all items (tx 0: mixed)
    // the schedule on the device, and the inputs it was planned for
    BatchSeg *segs_dev = nullptr;
    unsigned *off_dev = nullptr;
    size_t segs_cap = 0, off_cap = 0;
    unsigned slots = 0, nsegs = 0;
    std::vector<int> plan_key;
    std::string err;
    ~qms_batch() {
        if (s) hipStreamSynchronize(s);
        if (items_dev) hipFree(items_dev);
        if (segs_dev) hipFree(segs_dev);
        if (off_dev) hipFree(off_dev);
        if (s) hipStreamDestroy(s);
    }
};

namespace {

int bfail(qms_batch *b, const std::exception &e, int code) {
    if (b) b->err = e.what();
    g_last_error = e.what();
    return code;
}

static_assert(sizeof(PlanSeg) == sizeof(BatchSeg), "host and device segment layouts differ");

// Plans the launch for the designs' current tilings and grid size and
// uploads the schedule, unless it is the one already on the device.
void batch_plan(qms_batch *b) {
    std::vector<int> key{b->blocks};
    std::vector<PlanDesign> pd;
    for (const BatchItem &it : b->items) {
        const unsigned ncol = (unsigned)it.t.ntx * (unsigned)it.t.nty;
        pd.push_back(PlanDesign{ncol, it.g.nz, it.t.kz});
        key.insert(key.end(), {(int)ncol, it.g.nz, it.t.kz});
    }
    if (key == b->plan_key) return;
    b->plan_key.clear();           // no plan on the device until this one is uploaded
    const BatchPlan p = plan_batch(pd, (unsigned)b->blocks);
    if ((uint64_t)p.segs.size() >= 0xffffffffull) throw std::runtime_error("batch too large");
    if (p.segs.size() > b->segs_cap) {
        if (b->segs_dev) { check(hipStreamSynchronize(b->s), "sync"); check(hipFree(b->segs_dev), "hipFree"); }
        b->segs_dev = nullptr;
        b->segs_cap = 0;
        check(hipMalloc(&b->segs_dev, sizeof(BatchSeg) * p.segs.size()), "hipMalloc");
        b->segs_cap = p.segs.size();
    }
    if (p.off.size() > b->off_cap) {
        if (b->off_dev) { check(hipStreamSynchronize(b->s), "sync"); check(hipFree(b->off_dev), "hipFree"); }
        b->off_dev = nullptr;
        b->off_cap = 0;
        check(hipMalloc(&b->off_dev, sizeof(unsigned) * p.off.size()), "hipMalloc");
        b->off_cap = p.off.size();
    }
    if (!p.segs.empty())
        check(hipMemcpyAsync(b->segs_dev, p.segs.data(), sizeof(BatchSeg) * p.segs.size(), hipMemcpyHostToDevice,
                             b->s), "h2d");
    check(hipMemcpyAsync(b->off_dev, p.off.data(), sizeof(unsigned) * p.off.size(), hipMemcpyHostToDevice, b->s),
          "h2d");
    check(hipStreamSynchronize(b->s), "sync");
    b->slots = p.slots;
    b->nsegs = (unsigned)p.segs.size();
    b->plan_key = key;
}

// Work items for rows at per-design velocities v[3d..3d+2]; transient rows
// use each design's own dt.  Waits for every handle's stream first.
void batch_items(qms_batch *b, const double *v, int transient) {
    b->tx = -1;
    for (size_t d = 0; d < b->hs.size(); ++d) {
        qms_handle *h = b->hs[d];
        check(hipStreamSynchronize(h->s), "sync");
        const int tr = (transient && h->cfg.dt > 0) ? 1 : 0;
        const OpParams op = make_op(h, v + 3 * d, h->cfg.dt, tr);
        BatchItem &it = b->items[d];
        unsigned ncol = 0;
        it.g = h->g;
        it.t = ztiling(h->g, h->zchunk, h->colw, ncol);
        it.c = op_coeffs(h);
        it.k = make_opk(h->g, op);
        it.x = h->x;
        it.y = h->w;
        const int icls = it.c.cls ? 1 : 0;
        if (b->tx == -1) { b->tx = it.t.tx; b->cls = icls; }
        else if (b->tx != it.t.tx || b->cls != icls) b->tx = 0;
    }
    check(hipMemcpyAsync(b->items_dev, b->items.data(), sizeof(BatchItem) * b->items.size(), hipMemcpyHostToDevice,
                         b->s), "h2d");
    check(hipStreamSynchronize(b->s), "sync");
    batch_plan(b);
}

void batch_launch(qms_batch *b, int mode, const double *v, int transient) {
    if (mode == 0) {
        if (b->plan_key.empty() || !b->segs_dev || !b->off_dev) throw std::runtime_error("batch: no schedule on the device");
        if (b->blocks == 0)   // one workgroup per segment
            launch_operator_batch(b->items_dev, b->segs_dev, nullptr, b->nsegs, b->tx, b->cls, b->s);
        else
            launch_operator_batch(b->items_dev, b->segs_dev, b->off_dev, b->slots, b->tx, b->cls, b->s);
    } else {
        for (size_t d = 0; d < b->hs.size(); ++d) {
            qms_handle *h = b->hs[d];
            const int tr = (transient && h->cfg.dt > 0) ? 1 : 0;
            launch_apply(h->g, op_coeffs(h), make_op(h, v + 3 * d, h->cfg.dt, tr), h->x, h->w, b->s);
        }
    }
}


}  // namespace

extern "C" {

int qms_batch_create(qms_handle *const *handles, int n, qms_batch **out) {
    if (!handles || n < 1 || !out) { g_last_error = "qms_batch_create: no handles"; return QMS_ERR_ARG; }
    *out = nullptr;
    std::unique_ptr<qms_batch> b(new qms_batch());
    try {
        for (int d = 0; d < n; ++d) {
            if (!handles[d]) throw std::runtime_error("qms_batch_create: null handle");
            for (int e = 0; e < d; ++e)
                if (handles[e] == handles[d]) throw std::runtime_error("qms_batch_create: handle listed twice");
            b->hs.push_back(handles[d]);
        }
        b->items.resize(n);
        check(hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking), "hipStreamCreate");
        check(hipMalloc(&b->items_dev, sizeof(BatchItem) * n), "hipMalloc");
    } catch (const HipError &e) {
        return bfail(nullptr, e, e.code);
    } catch (const std::exception &e) {
        return bfail(nullptr, e, QMS_ERR_ARG);
    }
    *out = b.release();
    return QMS_OK;
}

int qms_batch_set_blocks(qms_batch *b, int blocks) {
    if (!b || blocks < 0) return QMS_ERR_ARG;
    b->blocks = blocks;
    return QMS_OK;
}

int qms_batch_apply(qms_batch *b, const double *v, int transient) {
    if (!b || !v) return QMS_ERR_ARG;
    try {
        batch_items(b, v, transient);
        batch_launch(b, 0, v, transient);
        check(hipStreamSynchronize(b->s), "sync");
    } catch (const HipError &e) {
        return bfail(b, e, e.code);
    } catch (const std::exception &e) {
        return bfail(b, e, QMS_ERR_ARG);
    }
    return QMS_OK;
}

int qms_batch_tune(qms_batch *b, const double *v, int transient, int reps, int *zchunks, int *colwidths,
                   int *blocks) {
    if (!b || !v || reps < 1) return QMS_ERR_ARG;
    try {
        static const int KZ[] = {1, 2, 3, 4, 6, 8};
        static const int CW[] = {64, 32};
        for (size_t d = 0; d < b->hs.size(); ++d) {
            qms_handle *h = b->hs[d];
            check(hipStreamSynchronize(h->s), "sync");
            const int tr = (transient && h->cfg.dt > 0) ? 1 : 0;
            double best = 1e30;
            int bestkz = h->zchunk, bestcw = h->colw;
            for (int cw : CW) {
                for (int kz : KZ) {
                    h->zchunk = kz;
                    h->colw = cw;
                    const OpParams op = make_op(h, v + 3 * d, h->cfg.dt, tr);
                    const double ms =
                        time_on(b->s, reps, [&]() { launch_apply(h->g, op_coeffs(h), op, h->x, h->w, b->s); });
                    if (ms < best) { best = ms; bestkz = kz; bestcw = cw; }
                }
            }
            h->zchunk = bestkz;
            h->colw = bestcw;
        }
        // The per-design choice favours many small columns (a lone small
        // design must fill the GPU); in a batch the designs fill it together,
        // so also try one shape for all designs, each with one workgroup per
        // natural piece and with the balanced persistent schedule at one and
        // two times the resident workgroup count, and keep the fastest.
        std::vector<int> bkz(b->hs.size()), bcw(b->hs.size());
        for (size_t d = 0; d < b->hs.size(); ++d) { bkz[d] = b->hs[d]->zchunk; bcw[d] = b->hs[d]->colw; }
        double best = 1e30;
        int ucw = 0, ukz = 0, bestb = 0;
        auto trial = [&](int cw, int kz) {
            for (size_t d = 0; d < b->hs.size(); ++d) {
                b->hs[d]->zchunk = kz ? kz : bkz[d];
                b->hs[d]->colw = kz ? cw : bcw[d];
            }
            b->blocks = 0;
            batch_items(b, v, transient);
            const int res = batch_resident_blocks(b->tx, b->cls);
            for (int bl : {0, res, 2 * res}) {
                b->blocks = bl;
                batch_plan(b);
                const double ms = time_on(b->s, reps, [&]() { batch_launch(b, 0, v, transient); });
                if (ms < best) { best = ms; ucw = cw; ukz = kz; bestb = bl; }
            }
        };
        trial(0, 0);
        for (int cw : CW)
            for (int kz : {2, 4, 8, 16, 32, 64}) trial(cw, kz);
        for (size_t d = 0; d < b->hs.size(); ++d) {
            b->hs[d]->zchunk = ukz ? ukz : bkz[d];
            b->hs[d]->colw = ukz ? ucw : bcw[d];
            if (zchunks) zchunks[d] = b->hs[d]->zchunk;
            if (colwidths) colwidths[d] = b->hs[d]->colw;
        }
        b->blocks = bestb;
        batch_items(b, v, transient);
        if (blocks) *blocks = bestb;
    } catch (const HipError &e) {
        return bfail(b, e, e.code);
    } catch (const std::exception &e) {
        return bfail(b, e, QMS_ERR_ARG);
    }
    return QMS_OK;
}

int qms_batch_time(qms_batch *b, int mode, const double *v, int transient, int reps, double *ms_per_launch,
                   double *bytes_per_launch) {
    if (!b || !v || reps < 1 || !ms_per_launch || !bytes_per_launch || mode < 0 || mode > 1) return QMS_ERR_ARG;
    try {
        batch_items(b, v, transient);
        *ms_per_launch = time_on(b->s, reps, [&]() { batch_launch(b, mode, v, transient); });
        double bytes = 0;
        for (qms_handle *h : b->hs) bytes += 8.0 * (2.0 * (double)h->n + 6.0 * (double)h->g.n1);
        *bytes_per_launch = bytes;
    } catch (const HipError &e) {
        return bfail(b, e, e.code);
    } catch (const std::exception &e) {
        return bfail(b, e, QMS_ERR_ARG);
    }
    return QMS_OK;
}

int qms_batch_plan(int n, const int *ncols, const int *nz, const int *band, int blocks, unsigned *segs,
                   int seg_cap, unsigned *slot_off, int off_cap, int *nsegs, int *nslots) {
    if (n < 1 || !ncols || !nz || !band || blocks < 0 || !nsegs || !nslots) {
        g_last_error = "qms_batch_plan: invalid argument";
        return QMS_ERR_ARG;
    }
    try {
        std::vector<PlanDesign> pd;
        for (int d = 0; d < n; ++d) {
            if (ncols[d] < 1 || nz[d] < 1 || band[d] < 1) {
                g_last_error = "qms_batch_plan: columns, planes and z-chunk must be positive";
                return QMS_ERR_ARG;
            }
            pd.push_back(PlanDesign{(unsigned)ncols[d], nz[d], band[d]});
        }
        const BatchPlan p = plan_batch(pd, (unsigned)blocks);
        if (p.segs.size() > 0x7fffffffu) { g_last_error = "qms_batch_plan: too many segments"; return QMS_ERR_ARG; }
        *nsegs = (int)p.segs.size();
        *nslots = (int)p.slots;
        if (!segs || !slot_off) return QMS_OK;
        if ((size_t)seg_cap < p.segs.size() || (size_t)off_cap < p.off.size()) {
            g_last_error = "qms_batch_plan: output buffers too small";
            return QMS_ERR_ARG;
        }
        for (size_t e = 0; e < p.segs.size(); ++e) {
            segs[4 * e] = p.segs[e].d;
            segs[4 * e + 1] = p.segs[e].col;
            segs[4 * e + 2] = (unsigned)p.segs[e].kb;
            segs[4 * e + 3] = (unsigned)p.segs[e].ke;
        }
        for (size_t s = 0; s < p.off.size(); ++s) slot_off[s] = p.off[s];
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return QMS_ERR_ARG;
    }
    return QMS_OK;
}

const char *qms_batch_last_error(const qms_batch *b) { return b ? b->err.c_str() : g_last_error.c_str(); }

void qms_batch_destroy(qms_batch *b) { delete b; }

void qms_destroy(qms_handle *h) { delete h; }

const char *qms_last_error(const qms_handle *h) {
    if (h) return h->err.c_str();
    return g_last_error.c_str();
}

}  // extern "C"
