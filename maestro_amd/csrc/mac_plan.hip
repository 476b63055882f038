// mac_plan.hip -- C ABI of libmaestro_mac.so: plans, macproject, solver
// driver (V-cycle, FP32-in-FP64 refinement) and the adjoint.
//
// Reference interfaces replaced:
//   macproject      Source/macproject.f90:51-328
//   mac_multigrid   Source/mac_multigrid.f90:18-71 (-> FBoxLib ml_cc_solve)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <numeric>
#include <thread>
#include <vector>

#include "mac_launch.h"

using namespace mac;

#define HIPCHK_V(x) (void)(x)
#define HIPCHK(x)                                                                     \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess) {                                                       \
            if (getenv("MAC_DEBUG"))                                                  \
                fprintf(stderr, "maestro_mac: %s failed: %s (%s:%d)\n", #x,           \
                        hipGetErrorString(e_), __FILE__, __LINE__);                   \
            return MAC_ERR_HIP;                                                       \
        }                                                                             \
    } while (0)

namespace {

template <typename T>
struct Hier {
    int nlev = 0, lc = 0;
    std::vector<Lev<T>> lev;
    Lev<T> *lev_dev = nullptr;
    std::vector<void *> allocs;
    bool built = false;
    std::vector<T *> alpha;   // per-level alpha storage (mac_solve), in the level's allocation
    bool alpha_on = false;    // whether lev[].alpha (and lev_dev) point at it
    // FP32 3D level 0 only: the rho form of the finest level (parts of its
    // allocation, so the row kernels reach them through the weights' buffer)
    RhoLev rho0 = {};
};

// A few host threads that run one parallel copy at a time: run(fn) calls
// fn(t) for t = 0 .. size()-1 (t = 0 on the calling thread) and returns when
// every call has returned.  One run at a time, so every worker sees each run.
class CopyPool {
  public:
    explicit CopyPool(int n) : n_(n)
    {
        for (int t = 1; t < n_; ++t) th_.emplace_back([this, t] { loop(t); });
    }
    ~CopyPool()
    {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
            ++gen_;
        }
        cv_.notify_all();
        for (auto &x : th_) x.join();
    }
    int size() const { return n_; }
    void run(const std::function<void(int)> &fn)
    {
        {
            std::lock_guard<std::mutex> g(m_);
            fn_ = &fn;
            left_ = n_ - 1;
            ++gen_;
        }
        cv_.notify_all();
        fn(0);
        std::unique_lock<std::mutex> g(m_);
        done_.wait(g, [this] { return left_ == 0; });
        fn_ = nullptr;
    }

  private:
    void loop(int t)
    {
        long seen = 0;
        for (;;) {
            const std::function<void(int)> *fn;
            {
                std::unique_lock<std::mutex> g(m_);
                cv_.wait(g, [&] { return gen_ != seen; });
                seen = gen_;
                if (stop_) return;
                fn = fn_;
            }
            (*fn)(t);
            std::lock_guard<std::mutex> g(m_);
            if (--left_ == 0) done_.notify_one();
        }
    }
    int n_;
    std::vector<std::thread> th_;
    std::mutex m_;
    std::condition_variable cv_, done_;
    const std::function<void(int)> *fn_ = nullptr;
    long gen_ = 0;
    int left_ = 0;
    bool stop_ = false;
};

struct HostStage {
    // device mirrors of the caller's per-box host arrays
    std::vector<double *> umac, phi, rho, rhs, cart;
    std::vector<size_t> umac_n, cell_n_phi, cell_n_rho, cell_n_rhs, cart_n;
    bool rhs_ready = false, cart_ready = false, base_ready = false;
    // pinned staging of the copies: two slots; the host threads fill (or
    // drain) one while the copy engine moves the other
    char *slot[2] = {nullptr, nullptr};
    hipEvent_t moved[2] = {nullptr, nullptr};
    CopyPool *pool = nullptr;
};

}  // namespace

struct mac_plan {
    int dim = 3, nboxes = 0;
    int n[3] = {1, 1, 1}, dlo[3] = {0, 0, 0};
    double dx[3] = {1, 1, 1};
    EllBC bc;
    int phys[3][2];
    bool singular = true;
    bool helm = false;            // inside mac_solve: alpha and the call's BCs active
    bool planar = false;          // the current projection has a planar beta0 (the rho form may apply)
    int ng_u = 0, ng_phi = 0, ng_rho = 0, ng_rhs = 0, ng_cart = 0;
    mac_opts opts;
    std::vector<BoxDev> boxes;
    BoxDev *boxes_dev = nullptr;
    int maxcells = 0, maxfaces = 0, maxghost = 0;
    int maxcols = 0, maxez = 1;   // largest x-y cross-section and z extent of a box (column kernels)
    long maxplane = 0;
    int *boxmap_dev = nullptr;
    int mapg[3] = {1, 1, 1}, mapn[3] = {1, 1, 1};
    Hier<double> H64;
    Hier<float> H32;
    // level 0 of H32 in compact storage (MAC_PREC_MIXED_BF16 where it applies)
    Compact16 c16 = {};
    std::vector<void *> c16_allocs;
    bool launch_failed = false;   // a compact-storage launch found no kernel
    double *part = nullptr;
    long part_cap = 0;
    double *mean_dev = nullptr;
    unsigned long long *nrm_dev = nullptr;
    unsigned long long *nrm_host = nullptr;   // pinned: [0] a copied norm, [1] a copied flag, [2] max |rh|
    // residual norms reach the host through mapped pinned memory: a
    // one-thread kernel writes the value and then a sequence number, which
    // the host polls (no stream synchronisation per refinement step)
    unsigned long long *nrm_map = nullptr;    // host pointer ([0] value, [1] sequence)
    unsigned long long *nrm_map_dev = nullptr;
    unsigned long long nrm_seq = 0;
    bool nrm_map_ok = true;
    // the projection's device flag of non-finite inputs: read with the
    // initial residual norm rather than by a stream synchronisation of its
    // own between the setup and the solve
    const int *setup_err = nullptr;
    int *err_dev = nullptr;
    double *cg = nullptr;
    long cg_cap = 0;
    double **tab_dev = nullptr;   // pointer tables
    std::vector<double *> tab_host;   // host side of a call's pointer table (sized at creation)
    double *b0_dev = nullptr, *b0e_dev = nullptr;
    int nr = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[8];
    HostStage hs;
    double *scratch_cells = nullptr;   // one FP64 cell array (adjoint, make_s0mac)
    void *call_tab = nullptr;          // device pointer tables of the producer calls
    size_t call_tab_cap = 0;
    int device = 0;
    // kernel timing (mac_plan_profile)
    bool prof_on = false;
    std::vector<hipEvent_t> prof_ev;   // pairs
    std::vector<int> prof_cls;
    int prof_used = 0;
    long prof_cnt[MAC_PROF_NCLASS] = {0};
    double prof_ms[MAC_PROF_NCLASS] = {0};
};

static constexpr int PROF_PAIRS = 512;

static void prof_flush(mac_plan *P)
{
    if (!P->prof_used) return;
    (void)hipStreamSynchronize(P->stream);
    for (int q = 0; q < P->prof_used; ++q) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, P->prof_ev[2 * q], P->prof_ev[2 * q + 1]) == hipSuccess) {
            P->prof_ms[P->prof_cls[q]] += ms;
            P->prof_cnt[P->prof_cls[q]] += 1;
        }
    }
    P->prof_used = 0;
}

// returns the slot to close with prof_end, or -1 when profiling is off
static int prof_begin(mac_plan *P, int cls)
{
    if (!P->prof_on) return -1;
    if (P->prof_used == PROF_PAIRS) prof_flush(P);
    int q = P->prof_used++;
    P->prof_cls[q] = cls;
    (void)hipEventRecord(P->prof_ev[2 * q], P->stream);
    return q;
}

static void prof_end(mac_plan *P, int q)
{
    if (q >= 0) (void)hipEventRecord(P->prof_ev[2 * q + 1], P->stream);
}

static long ncell(const int *n) { return (long)n[0] * n[1] * n[2]; }

static int ell_kind(int phys)
{
    if (phys == MAC_PERIODIC) return ELL_PER;
    if (phys == MAC_OUTLET) return ELL_DIR;
    return ELL_NEU;
}

extern "C" int mac_default_opts(mac_opts *o)
{
    if (!o) return MAC_ERR_INVALID;
    memset(o, 0, sizeof(*o));
    o->eps_rel = 1e-10;
    o->eps_abs = -1.0;
    o->max_iter = 200;
    o->nu1 = 2;
    o->nu2 = 2;
    o->nub = 10;
    o->bottom_eps = 1e-3;
    o->bottom_max_iter = 1000;
    o->bottom_solver = MAC_BOTTOM_CG;
    o->max_levels = 0;
    o->use_alt_energy_fix = 1;
    o->precision = MAC_PREC_FP64;
    o->inner_cycles = 0;
    o->device_ptrs = 0;
    o->verbose = 0;
    return MAC_OK;
}

extern "C" const char *mac_version(void) { return "maestro_mac 0.1 (gfx950)"; }

// ------------------------------------------------------------ hierarchy

// Levels of at most this many cells run inside the single-workgroup coarse
// cycle (8^3 in 3D, 32^2 in 2D).  A 16^3 level fits its LDS in FP32 but is
// slower there than as separate multi-block launches: one workgroup runs on
// one CU.
static int coarse_threshold(int dim) { return dim == 3 ? 512 : 1024; }

template <typename T>
static int build_hier(mac_plan *P, Hier<T> &H)
{
    if (H.built) return MAC_OK;
    int n[3] = {P->n[0], P->n[1], P->n[2]};
    double h[3] = {P->dx[0], P->dx[1], P->dx[2]};
    int nl = 1;
    {
        int m[3] = {n[0], n[1], n[2]};
        for (;;) {
            if (P->opts.max_levels > 0 && nl >= P->opts.max_levels) break;
            bool ok = true;
            for (int d = 0; d < P->dim; ++d)
                if ((m[d] & 1) || m[d] / 2 < 2) ok = false;
            if (!ok) break;
            for (int d = 0; d < P->dim; ++d) m[d] /= 2;
            ++nl;
        }
    }
    H.nlev = nl;
    H.lev.resize(nl);
    for (int l = 0; l < nl; ++l) {
        Lev<T> &L = H.lev[l];
        L.nx = n[0]; L.ny = n[1]; L.nz = n[2];
        L.perx = P->bc.k[0][0] == ELL_PER;
        L.pery = P->bc.k[1][0] == ELL_PER;
        L.perz = (P->dim == 3) && P->bc.k[2][0] == ELL_PER;
        L.ih2x = (T)(1.0 / (h[0] * h[0]));
        L.ih2y = (T)(1.0 / (h[1] * h[1]));
        L.ih2z = (P->dim == 3) ? (T)(1.0 / (h[2] * h[2])) : T(0);
        L.omega = T(1);
        // the level's arrays share one allocation (256-byte aligned parts), so
        // a kernel can reach all of them through one buffer resource
        const long nc = ncell(n);
        // the last part holds alpha (mac_solve); L.alpha points at it only
        // while a solve with alpha runs
        T *alpha_part = nullptr;
        // the FP32 3D finest level also holds its rho form: rho, the plane and
        // z-face factors and the low z slab
        T *rho_p = nullptr, *cxy_p = nullptr, *cz_p = nullptr, *bzl_p = nullptr;
        T **arr[15] = {&L.phi, &L.phi2, &L.rhs, &L.res, &L.bx, &L.by, &L.bz, &L.bxh, &L.byh, &L.bzh, &alpha_part,
                       &rho_p, &cxy_p, &cz_p, &bzl_p};
        // second iterate: ping-pong of the fused 3D sweep, and on level 0
        // the output of the fused refinement update
        const bool has2 = P->dim == 3 || l == 0, d3 = P->dim == 3;
        const bool rh = sizeof(T) == 4 && l == 0 && d3;
        const long len[15] = {nc, has2 ? nc : 0, nc, nc, nc, nc, d3 ? nc : 0, (long)n[1] * n[2],
                              (long)n[0] * n[2], d3 ? (long)n[0] * n[1] : 0, nc,
                              rh ? nc : 0, rh ? (long)n[2] : 0, rh ? (long)n[2] + 1 : 0, rh ? (long)n[0] * n[1] : 0};
        auto part = [](long cnt) { return ((cnt * (long)sizeof(T) + 255) / 256) * 256; };
        long total = 0;
        for (int a = 0; a < 15; ++a) total += part(len[a]);
        void *blk = nullptr;
        HIPCHK(hipMalloc(&blk, (size_t)total));
        H.allocs.push_back(blk);
        long off = 0;
        for (int a = 0; a < 15; ++a) {
            *arr[a] = len[a] ? (T *)((char *)blk + off) : nullptr;
            off += part(len[a]);
        }
        if (rh) H.rho0 = RhoLev{(float *)rho_p, (float *)cxy_p, (float *)cz_p, (float *)bzl_p};
        L.alpha = nullptr;
        H.alpha.push_back(alpha_part);
        for (int d = 0; d < P->dim; ++d) { n[d] /= 2; h[d] *= 2.0; }
    }
    int thr = coarse_threshold(P->dim);
    H.lc = nl - 1;
    for (int l = 0; l < nl; ++l)
        if ((long)H.lev[l].nx * H.lev[l].ny * H.lev[l].nz <= thr) { H.lc = l; break; }
    void *q = nullptr;
    HIPCHK(hipMalloc(&q, sizeof(Lev<T>) * nl));
    H.allocs.push_back(q);
    H.lev_dev = (Lev<T> *)q;
    HIPCHK(hipMemcpy(H.lev_dev, H.lev.data(), sizeof(Lev<T>) * nl, hipMemcpyHostToDevice));
    // bottom solver scratch (doubles)
    const Lev<T> &B = H.lev[nl - 1];
    long nb = (long)B.nx * B.ny * B.nz;
    if (BOTTOM_VECS * nb > P->cg_cap) {
        if (P->cg) (void)hipFree(P->cg);
        HIPCHK(hipMalloc(&P->cg, sizeof(double) * BOTTOM_VECS * nb));
        P->cg_cap = BOTTOM_VECS * nb;
    }
    H.built = true;
    return MAC_OK;
}

template <typename T>
static void free_hier(Hier<T> &H)
{
    for (void *p : H.allocs) (void)hipFree(p);
    H.allocs.clear();
    H.lev.clear();
    H.alpha.clear();
    H.alpha_on = false;
    H.built = false;
}

// point the levels' alpha at the hierarchy's alpha storage (on) or at
// nothing (off); the device copy of the level table follows
template <typename T>
static int set_alpha(mac_plan *P, Hier<T> &H, bool on)
{
    if (!H.built || on == H.alpha_on) return MAC_OK;
    for (int l = 0; l < H.nlev; ++l) H.lev[l].alpha = on ? H.alpha[l] : nullptr;
    HIPCHK(hipMemcpyAsync(H.lev_dev, H.lev.data(), sizeof(Lev<T>) * H.nlev, hipMemcpyHostToDevice, P->stream));
    H.alpha_on = on;
    return MAC_OK;
}

template <typename T>
static void coarsen_all(mac_plan *P, Hier<T> &H)
{
    for (int l = 1; l < H.nlev; ++l) launch_coarsen<T>(P->dim, H.lev[l - 1], H.lev[l], P->stream);
}

template <typename T>
static void smooth(mac_plan *P, const Lev<T> &L, int level, int sweeps)
{
    const int cls = sizeof(T) == 8 ? MAC_PROF_GSRB64 : MAC_PROF_GSRB32;
    for (int it = 0; it < sweeps; ++it)
        for (int color = 0; color < 2; ++color) {
            int q = (level == 0) ? prof_begin(P, cls) : -1;
            launch_gsrb<T>(P->dim, L, color, P->stream);
            prof_end(P, q);
        }
}

// 3D levels run the fused red-black sweeps (pin -> pout, one launch per
// sweep); 2D levels sweep per colour in place
static bool use_fused(const mac_plan *P) { return P->dim == 3; }

// K: level l's weights and right side in compact storage (level 0 only)
// nsweep fused sweeps of a 3D level; the iterate moves between phi and phi2
// (cur tracks which holds it).  first_mode applies to the first sweep.
template <typename T>
static void fused_sweeps(mac_plan *P, Hier<T> &H, int l, int nsweep, int first_mode, int &cur,
                         const Compact16 *K = nullptr, const RhoLev *Rh = nullptr)
{
    const bool f64 = sizeof(T) == 8;
    Lev<T> &L = H.lev[l];
    for (int it = 0; it < nsweep; ++it) {
        T *pin = cur ? L.phi2 : L.phi;
        T *pout = cur ? L.phi : L.phi2;
        int mode = it == 0 ? first_mode : SWEEP_PLAIN;
        const int cls = mode == SWEEP_ZERO ? (f64 ? MAC_PROF_GSRB64_ZERO : MAC_PROF_GSRB32_ZERO)
                                           : (f64 ? MAC_PROF_GSRB64 : MAC_PROF_GSRB32);
        int q = (l == 0) ? prof_begin(P, cls) : -1;
        bool ok;
        if constexpr (sizeof(T) == 4)
            ok = Rh ? launch_gsrb_rho(L, *Rh, pin, pout, mode, P->stream)
                    : launch_gsrb_fused<T>(L, pin, pout, mode, P->stream, K);
        else
            ok = launch_gsrb_fused<T>(L, pin, pout, mode, P->stream, K);
        if (ok) {
            cur ^= 1;
        } else if (L.alpha && !K) {
            // a level with alpha that no fused sweep takes: the two colours in
            // place on the current iterate (same result as a fused sweep)
            Lev<T> Lc = L;
            Lc.phi = pin;
            if (mode == SWEEP_ZERO) launch_fill<T>(pin, (long)L.nx * L.ny * L.nz, T(0), P->stream);
            launch_gsrb<T>(3, Lc, 0, P->stream);
            launch_gsrb<T>(3, Lc, 1, P->stream);
        } else {
            P->launch_failed = true;
        }
        prof_end(P, q);
    }
}

// phi += P e on a 3D level's current iterate (the coarse correction as a
// separate streaming pass: interpolating inside the sweep stalls its plane
// pipeline on the coarse loads)
template <typename T>
static void prolong_cur(mac_plan *P, Hier<T> &H, int l, int cur)
{
    Lev<T> &L = H.lev[l];
    Lev<T> Lc = L;
    Lc.phi = cur ? L.phi2 : L.phi;
    int q = (l == 0) ? prof_begin(P, sizeof(T) == 8 ? MAC_PROF_PROLONG64 : MAC_PROF_PROLONG32) : -1;
    launch_prolong<T>(3, Lc, H.lev[l + 1], P->bc, P->stream);
    prof_end(P, q);
}

static CoarseOpts coarse_opts(const mac_plan *P, int type)
{
    const mac_opts &o = P->opts;
    CoarseOpts co;
    co.cycle = type;
    co.bottom = o.bottom_solver;
    co.nu1 = o.nu1; co.nu2 = o.nu2; co.nub = o.nub;
    co.bottom_eps = o.bottom_eps;
    co.bottom_maxit = o.bottom_max_iter;
    co.singular = P->singular;
    co.bc = P->bc;
    return co;
}

// The shape of the cycle's next visit of level l + 1 after its first one
// (0: none): a W visits twice (each visit a W), an F first as an F and then
// as a V, a V once.
static int second_visit(int type)
{
    return type == MAC_CYCLE_W ? MAC_CYCLE_W : type == MAC_CYCLE_F ? MAC_CYCLE_V : 0;
}

static int cycle_shape(const mac_opts &o) { return o.cycle_type == 0 ? MAC_CYCLE_V : o.cycle_type; }

// One cycle of shape `type` from level l (levels lc.. run inside the
// single-workgroup coarse cycle).  zero: level l's iterate is known to be
// zero (a first visit below level 0, or a refinement step at level 0); a
// second visit starts from the first one's iterate.  K: level 0's weights
// and right side in compact storage; Rh: level 0 in the rho form (fused path).
template <typename T>
static void cycle_from(mac_plan *P, Hier<T> &H, int l, int type, bool zero, const Compact16 *K,
                       const RhoLev *Rh = nullptr)
{
    hipStream_t s = P->stream;
    const int dim = P->dim;
    const mac_opts &o = P->opts;
    Lev<T> &L = H.lev[l];
    const long n = (long)L.nx * L.ny * L.nz;
    if (l == H.lc) {
        // below level 0 the fused restriction has zeroed the iterate already
        if (zero && (l == 0 || !use_fused(P))) launch_fill<T>(L.phi, n, T(0), s);
        int q = prof_begin(P, MAC_PROF_COARSE);
        launch_coarse_cycle<T>(dim, H.lev_dev, H.lev.data(), H.lc, H.nlev, coarse_opts(P, type), P->cg, s);
        prof_end(P, q);
        return;
    }
    const int again = second_visit(type);
    if (use_fused(P)) {
        int cur = 0;   // 1: the iterate is in phi2
        const Compact16 *Kl = l == 0 ? K : nullptr;
        const RhoLev *Rl = l == 0 ? Rh : nullptr;
        if (o.nu1 > 0) fused_sweeps<T>(P, H, l, o.nu1, zero ? SWEEP_ZERO : SWEEP_PLAIN, cur, Kl, Rl);
        else if (zero) launch_fill<T>(L.phi, n, T(0), s);
        int q = (l == 0) ? prof_begin(P, sizeof(T) == 8 ? MAC_PROF_RESTRICT64 : MAC_PROF_RESTRICT32) : -1;
        // also zeroes level l + 1's iterate
        if (!launch_resid_restrict<T>(dim, L, cur ? L.phi2 : L.phi, H.lev[l + 1], s, Kl, Rl))
            P->launch_failed = true;
        prof_end(P, q);
        cycle_from<T>(P, H, l + 1, type, true, K);
        if (again) cycle_from<T>(P, H, l + 1, again, false, K);
        prolong_cur<T>(P, H, l, cur);
        if (o.nu2 > 0) fused_sweeps<T>(P, H, l, o.nu2, SWEEP_PLAIN, cur, Kl, Rl);
        if (cur)   // bring the iterate back to phi
            HIPCHK_V(hipMemcpyAsync(L.phi, L.phi2, sizeof(T) * (size_t)n, hipMemcpyDeviceToDevice, s));
        return;
    }
    if (zero) launch_fill<T>(L.phi, n, T(0), s);
    smooth<T>(P, L, l, o.nu1);
    launch_residual<T>(dim, L, L.res, nullptr, s);
    launch_restrict<T>(dim, L.res, L.nx, L.ny, H.lev[l + 1], s);
    cycle_from<T>(P, H, l + 1, type, true, K);
    if (again) cycle_from<T>(P, H, l + 1, again, false, K);
    launch_prolong<T>(dim, L, H.lev[l + 1], P->bc, s);
    smooth<T>(P, L, l, o.nu2);
}

// One cycle of the plan's shape on level 0.  zero_first: the level-0
// iterate is known to be zero (refinement step); K: level 0's weights and
// right side in compact storage (fused path).
template <typename T>
static void vcycle(mac_plan *P, Hier<T> &H, bool zero_first = false, const Compact16 *K = nullptr,
                   const RhoLev *Rh = nullptr)
{
    cycle_from<T>(P, H, 0, cycle_shape(P->opts), zero_first, K, Rh);
}

// whether level 0 of the FP32 hierarchy is kept in compact storage
// (MAC_PREC_MIXED_BF16 where every level-0 kernel of the mixed path has its
// compact form)
static bool compact_applies(mac_plan *P)
{
    return P->opts.precision == MAC_PREC_MIXED_BF16 && P->H32.built && use_fused(P) && !P->H64.alpha_on &&
           compact_level_ok(P->dim, P->H32.lev[0]);
}

// whether the FP32 finest level runs in the rho form: MAC_PREC_MIXED, a 3D
// projection with a planar beta0 (the face coefficients are 2 beta0^p /
// (rho_L + rho_R)), x and y periodic, nx = 128 or 256, at least two levels
static bool rho_mode(mac_plan *P)
{
    return P->opts.precision == MAC_PREC_MIXED && P->dim == 3 && P->planar && !P->helm && !P->H64.alpha_on &&
           P->H32.built && P->H32.nlev > 1 && P->H32.rho0.rho && rho_level_ok(3, P->H32.lev[0]);
}

// Level 0 of the FP32 hierarchy in compact storage: *K is set where it
// applies, else left null and the FP32 arrays are used.  from_f64: the
// weights come from the FP64 level 0 in one pass that also writes the FP32
// level-1 weights; otherwise from H32 level 0.  The right side is written by
// the refinement residual.
static int compact_level0(mac_plan *P, const Compact16 **K, bool from_f64)
{
    *K = nullptr;
    if (!compact_applies(P)) return MAC_OK;
    const Lev<float> &F = P->H32.lev[0];
    Compact16 &C = P->c16;
    if (P->c16_allocs.empty()) {   // one allocation (one buffer resource for the kernels)
        const long nc = (long)F.nx * F.ny * F.nz;
        unsigned short **arr[7] = {&C.bx, &C.by, &C.bz, &C.bxh, &C.byh, &C.bzh, &C.rhs};
        const long len[7] = {nc, nc, nc, (long)F.ny * F.nz, (long)F.nx * F.nz, (long)F.nx * F.ny, nc};
        auto part = [](long cnt) { return ((cnt * 2L + 255) / 256) * 256; };
        long total = 0;
        for (int a = 0; a < 7; ++a) total += part(len[a]);
        void *q = nullptr;
        if (hipMalloc(&q, (size_t)total) != hipSuccess) {
            C = Compact16{};
            return MAC_ERR_NOMEM;
        }
        P->c16_allocs.push_back(q);
        long off = 0;
        for (int a = 0; a < 7; ++a) {
            *arr[a] = (unsigned short *)((char *)q + off);
            off += part(len[a]);
        }
    }
    const long nc = (long)F.nx * F.ny * F.nz;
    if (from_f64 && P->H32.nlev > 1) {
        launch_compact_level0(P->H64.lev[0], C, P->H32.lev[1], P->stream);
        *K = &C;
        return MAC_OK;
    }
    launch_to_bf16(F.bx, C.bx, nc, P->stream);
    launch_to_bf16(F.by, C.by, nc, P->stream);
    launch_to_bf16(F.bz, C.bz, nc, P->stream);
    launch_to_bf16(F.bxh, C.bxh, (long)F.ny * F.nz, P->stream);
    launch_to_bf16(F.byh, C.byh, (long)F.nx * F.nz, P->stream);
    launch_to_bf16(F.bzh, C.bzh, (long)F.nx * F.ny, P->stream);
    *K = &C;
    return MAC_OK;
}

// Residual norm slot: nrm_dev[1], accumulated by the residual kernels
// (atomic max) and cleared again by the publishing kernel after each read.
static unsigned long long *res_slot(mac_plan *P) { return P->nrm_dev + 1; }

// Publish the residual norm accumulated in the slot and wait for it: a
// one-thread kernel copies it (also into nrm_dev[3]) with a sequence number
// into mapped host memory and clears the slot; the host polls the number.
// Without working mapped memory the value comes through a plain copy.  bad
// (where given) receives the projection's non-finite-input flag.
static int read_norm(mac_plan *P, double *out, int *bad = nullptr)
{
    unsigned long long u = 0, seq = 0;
    bool have = false;
    const int *flag = bad ? P->setup_err : nullptr;
    if (bad) *bad = 0;
    if (P->nrm_map_ok) {
        seq = ++P->nrm_seq;
        launch_publish_norm(res_slot(P), P->nrm_map_dev, seq, P->stream, P->nrm_dev + 3, flag);
        HIPCHK(hipGetLastError());
        volatile unsigned long long *h = P->nrm_map;
        for (long it = 0;; ++it) {
            if (h[1] == seq) { have = true; break; }
            if ((it & 1023) == 1023) {
                const hipError_t e = hipStreamQuery(P->stream);
                if (e == hipSuccess) {   // drained: the value is visible unless the mapping failed
                    std::atomic_thread_fence(std::memory_order_seq_cst);
                    if (h[1] == seq) have = true;
                    else P->nrm_map_ok = false;
                    break;
                }
                if (e != hipErrorNotReady) return MAC_ERR_HIP;
            }
        }
        std::atomic_thread_fence(std::memory_order_acquire);
        if (have) {
            u = h[0];
            if (flag) *bad = (int)h[2];
        }
    }
    if (!have) {   // plain copy: the published value's device copy, or the slot itself (then cleared)
        const unsigned long long *src = seq != 0 ? P->nrm_dev + 3 : res_slot(P);
        HIPCHK(hipMemcpyAsync(P->nrm_host, src, sizeof(unsigned long long), hipMemcpyDeviceToHost, P->stream));
        if (flag)
            HIPCHK(hipMemcpyAsync(P->nrm_host + 1, flag, sizeof(int), hipMemcpyDeviceToHost, P->stream));
        if (seq == 0) HIPCHK(hipMemsetAsync(res_slot(P), 0, sizeof(unsigned long long), P->stream));
        HIPCHK(hipStreamSynchronize(P->stream));
        u = *P->nrm_host;
        if (flag) memcpy(bad, P->nrm_host + 1, sizeof(int));
    }
    double v;
    memcpy(&v, &u, sizeof v);
    *out = v;
    return MAC_OK;
}

static int fp64_norm(mac_plan *P, double *out, int *bad = nullptr)
{
    const Lev<double> &L = P->H64.lev[0];
    int q = prof_begin(P, MAC_PROF_RESID64);
    launch_residual<double>(P->dim, L, (double *)nullptr, res_slot(P), P->stream);
    prof_end(P, q);
    return read_norm(P, out, bad);
}

static int ir_norm(mac_plan *P, double *out, const Compact16 *K, int *bad)
{
    int q = prof_begin(P, MAC_PROF_RESID64);
    launch_ir_residual(P->dim, P->H64.lev[0], P->H32.lev[0].rhs, res_slot(P), P->stream,
                       K ? K->rhs : nullptr);
    prof_end(P, q);
    return read_norm(P, out, bad);
}

static bool converged(const mac_opts &o, double r, double r0)
{
    if (r0 == 0.0) return true;
    if (r <= o.eps_rel * r0) return true;
    if (o.eps_abs > 0.0 && r <= o.eps_abs) return true;
    return false;
}

// FP32 V-cycles per refinement step; 0 picks by the storage level 0 actually
// uses (compact: whether the finest level is held in bfloat16).  Three
// cycles (error reduced by ~1e-3 per step at the V-cycle's rate) replace two
// FP64 update+residual passes of 56 B/cell each by FP32 passes inside the
// cycle; with bfloat16 storage the rounded operator limits a step's
// reduction to ~2^-9, so a third cycle buys nothing and two are used.
// Measured at 256^3 (9 V-cycles in all): FP32 storage 1/2/3 cycles per step
// 92.5/93.0/105.0 solves/s; bfloat16 storage 115.2/117.6/87.6 solves/s.
static int inner_per_step(const mac_opts &o, bool compact)
{
    if (o.inner_cycles > 0) return o.inner_cycles;
    return compact ? 2 : 3;
}

// Solve A phi = rhs on level 0 of H64 (phi holds the warm start).
// Over-relaxation of the red-black sweeps on the levels above the
// single-workgroup coarse cycle (which keeps Gauss-Seidel): mac_opts.omega
// where set, else 1 (FBoxLib's GSRB).  With 1.25 the 256^3 plane-parallel
// projection converges in 6 V-cycles instead of 9 and the 512^3 star in 12
// instead of 16 (to 1e-10, the oracle's FP64 cycle), but at the same relative
// residual the over-relaxed iterate leaves more of its error in smooth modes:
// 6.5e-9 from the 1e-13 solution in phi at 256^3 against 5.1e-10, so it is
// an option, not the default.
static double sweep_omega(const mac_plan *P)
{
    return P->opts.omega > 0 ? P->opts.omega : 1.0;
}

template <typename T>
static void set_omega(Hier<T> &H, double w)
{
    for (int l = 0; l < H.nlev; ++l) H.lev[l].omega = (l < H.lc) ? (T)w : T(1);
}

static int solve(mac_plan *P, mac_stats *st)
{
    int rc;
    const mac_opts &o = P->opts;
    if (o.precision < MAC_PREC_FP64 || o.precision > MAC_PREC_MIXED_BF16) return MAC_ERR_INVALID;
    if (o.cycle_type < 0 || o.cycle_type > MAC_CYCLE_V) return MAC_ERR_INVALID;
    if (o.bottom_solver < MAC_BOTTOM_SMOOTH || o.bottom_solver > MAC_BOTTOM_CG) return MAC_ERR_INVALID;
    if (!(o.omega >= 0 && o.omega < 2)) return MAC_ERR_INVALID;
    double r0 = 0, r = 0;
    int it = 0, inner = 0;
    HIPCHK(hipMemsetAsync(res_slot(P), 0, sizeof(unsigned long long), P->stream));
    if (o.precision == MAC_PREC_MIXED || o.precision == MAC_PREC_MIXED_BF16) {
        if ((rc = build_hier<float>(P, P->H32))) return rc;
        set_omega(P->H32, sweep_omega(P));
        // alpha in FP32 (from the FP64 level 0 and by coarsening) where the FP64 level has it
        if ((rc = set_alpha<float>(P, P->H32, P->helm && P->H64.alpha_on))) return rc;
        const Compact16 *K = nullptr;
        const RhoLev *Rh = nullptr;
        if (rho_mode(P)) {
            // the finest level in the rho form (rho written by the setup):
            // its face factors and z slabs, and FP32 level 1 from the FP64
            // weights, in one pass; then the rest
            launch_rho_prep(P->H64.lev[0], P->H32.rho0, P->H32.lev[0], P->H32.lev[1], P->b0_dev, P->b0e_dev,
                            P->opts.use_alt_energy_fix, P->stream);
            for (int l = 2; l < P->H32.nlev; ++l)
                launch_coarsen<float>(P->dim, P->H32.lev[l - 1], P->H32.lev[l], P->stream);
            Rh = &P->H32.rho0;
        } else if (compact_applies(P) && P->H32.nlev > 1) {
            // compact level 0 and FP32 level 1 in one pass, then the rest
            if ((rc = compact_level0(P, &K, true))) return rc;
            for (int l = 2; l < P->H32.nlev; ++l)
                launch_coarsen<float>(P->dim, P->H32.lev[l - 1], P->H32.lev[l], P->stream);
        } else {
            launch_weights_to_f32(P->dim, P->H64.lev[0], P->H32.lev[0], P->stream);
            coarsen_all<float>(P, P->H32);
            if ((rc = compact_level0(P, &K, false))) return rc;
        }
        P->launch_failed = false;
        int bad = 0;
        if ((rc = ir_norm(P, &r0, K, &bad))) return rc;
        if (bad) return MAC_ERR_NONFINITE;
        r = r0;
        const Lev<float> &E = P->H32.lev[0];
        const int per_step = inner_per_step(o, K != nullptr);
        Lev<double> &L0 = P->H64.lev[0];
        bool swapped = false;
        while (rc == MAC_OK && !converged(o, r, r0) && it < o.max_iter) {
            for (int c = 0; c < per_step; ++c) {
                vcycle<float>(P, P->H32, c == 0, K, Rh);
                ++inner;
            }
            // phi64 + e32 lands in phi2 together with the next residual;
            // the two buffers then trade names
            int q = prof_begin(P, MAC_PROF_IR_UPDATE);
            launch_ir_update_residual(P->dim, L0, E.phi, L0.phi2, E.rhs, res_slot(P), P->stream,
                                      K ? K->rhs : nullptr);
            prof_end(P, q);
            std::swap(L0.phi, L0.phi2);
            swapped = !swapped;
            if ((rc = read_norm(P, &r))) break;
            if (P->launch_failed) { rc = MAC_ERR_INVALID; break; }
            ++it;
            if (o.verbose) fprintf(stderr, "maestro_mac: IR step %d  res/res0 = %.3e\n", it, r / r0);
        }
        // the device copy of the level table follows the buffer swaps, on
        // every exit of the loop (a later coarse cycle may read level 0)
        if (swapped)
            HIPCHK(hipMemcpyAsync(P->H64.lev_dev, &L0, sizeof(Lev<double>), hipMemcpyHostToDevice, P->stream));
        if (rc) return rc;
    } else {
        set_omega(P->H64, sweep_omega(P));
        // the FP64 hierarchy's coarse weights (the mixed paths coarsen the FP32 one)
        coarsen_all<double>(P, P->H64);
        int bad = 0;
        if ((rc = fp64_norm(P, &r0, &bad))) return rc;
        if (bad) return MAC_ERR_NONFINITE;
        r = r0;
        while (!converged(o, r, r0) && it < o.max_iter) {
            vcycle<double>(P, P->H64);
            if ((rc = fp64_norm(P, &r))) return rc;
            ++it;
            if (o.verbose) fprintf(stderr, "maestro_mac: cycle %d  res/res0 = %.3e\n", it, r / r0);
        }
    }
    if (st) {
        st->iters = it;
        st->inner_cycles = inner;
        st->res0 = r0;
        st->res = r;
        st->converged = converged(o, r, r0);
        st->nlevels = P->H64.nlev;
    }
    return converged(o, r, r0) ? MAC_OK : MAC_ERR_NOT_CONVERGED;
}

static int ensure_part(mac_plan *P, long nparts)
{
    if (nparts > P->part_cap) {
        if (P->part) (void)hipFree(P->part);
        HIPCHK(hipMalloc(&P->part, sizeof(double) * 2 * nparts));
        P->part_cap = nparts;
    }
    return MAC_OK;
}

// subtract the compensated mean of x (n cells) on the plan stream
static int remove_mean(mac_plan *P, double *x, long n)
{
    long np = (n + 255) / 256;
    int rc;
    if ((rc = ensure_part(P, np))) return rc;
    launch_partial_sum<double>(x, n, P->part, P->stream);
    launch_mean(P->part, np, (double)n, P->mean_dev, P->stream);
    launch_sub_mean<double>(x, n, P->mean_dev, P->stream);
    return MAC_OK;
}

// ---------------------------------------------------------------- plan
// allow_holes: the boxes need not cover the domain (the owner map keeps -1
// where none does); for the fine level of a two-level projection only
static int plan_create(const mac_layout *lay, const mac_opts *opts, mac_plan **out, bool allow_holes);

extern "C" int mac_plan_create(const mac_layout *lay, const mac_opts *opts, mac_plan **out)
{
    return plan_create(lay, opts, out, false);
}

static int plan_create(const mac_layout *lay, const mac_opts *opts, mac_plan **out, bool allow_holes)
{
    if (!lay || !out) return MAC_ERR_INVALID;
    *out = nullptr;
    if (lay->dim != 2 && lay->dim != 3) return MAC_ERR_INVALID;
    if (lay->nboxes <= 0 || !lay->box_lo || !lay->box_hi) return MAC_ERR_LAYOUT;
    if (lay->ng_rho < 1 || lay->ng_umac < 0 || lay->ng_phi < 0 || lay->ng_macrhs < 0 ||
        lay->ng_beta0_cart < 0)
        return MAC_ERR_LAYOUT;
    mac_plan *P = new mac_plan();
    P->dim = lay->dim;
    P->nboxes = lay->nboxes;
    if (opts) P->opts = *opts;
    else mac_default_opts(&P->opts);
    for (int d = 0; d < 3; ++d) {
        if (d < lay->dim) {
            P->dlo[d] = lay->domain_lo[d];
            P->n[d] = lay->domain_hi[d] - lay->domain_lo[d] + 1;
            P->dx[d] = lay->dx[d];
            for (int s = 0; s < 2; ++s) {
                P->phys[d][s] = lay->phys_bc[d][s];
                P->bc.k[d][s] = ell_kind(lay->phys_bc[d][s]);
            }
            if ((P->bc.k[d][0] == ELL_PER) != (P->bc.k[d][1] == ELL_PER)) { delete P; return MAC_ERR_INVALID; }
            if (P->n[d] < 2 || !(P->dx[d] > 0)) { delete P; return MAC_ERR_LAYOUT; }
            // red-black colouring needs an even period across a periodic seam
            if (P->bc.k[d][0] == ELL_PER && (P->n[d] & 1)) { delete P; return MAC_ERR_LAYOUT; }
        } else {
            P->dlo[d] = 0; P->n[d] = 1; P->dx[d] = 1.0;
            P->phys[d][0] = P->phys[d][1] = MAC_PERIODIC;
            P->bc.k[d][0] = P->bc.k[d][1] = ELL_PER;
        }
    }
    P->singular = true;
    for (int d = 0; d < P->dim; ++d)
        for (int s = 0; s < 2; ++s)
            if (P->bc.k[d][s] == ELL_DIR) P->singular = false;
    P->ng_u = lay->ng_umac; P->ng_phi = lay->ng_phi; P->ng_rho = lay->ng_rho;
    P->ng_rhs = lay->ng_macrhs; P->ng_cart = lay->ng_beta0_cart;
    // boxes: inside the domain, disjoint, covering it exactly
    long covered = 0;
    P->boxes.resize(P->nboxes);
    int g[3] = {0, 0, 0};
    for (int b = 0; b < P->nboxes; ++b) {
        BoxDev B;
        for (int d = 0; d < 3; ++d) {
            B.lo[d] = (d < P->dim) ? lay->box_lo[3 * b + d] : 0;
            B.hi[d] = (d < P->dim) ? lay->box_hi[3 * b + d] : 0;
        }
        long v = 1;
        for (int d = 0; d < P->dim; ++d) {
            if (B.lo[d] < P->dlo[d] || B.hi[d] > P->dlo[d] + P->n[d] - 1 || B.hi[d] < B.lo[d]) {
                delete P;
                return MAC_ERR_LAYOUT;
            }
            v *= (B.hi[d] - B.lo[d] + 1);
            g[d] = std::gcd(g[d], std::gcd(B.lo[d] - P->dlo[d], B.hi[d] - B.lo[d] + 1));
        }
        covered += v;
        P->boxes[b] = B;
    }
    if (covered != ncell(P->n) && !(allow_holes && covered < ncell(P->n))) { delete P; return MAC_ERR_LAYOUT; }
    // owner map at the gcd resolution; also detects overlaps
    for (int d = 0; d < 3; ++d) {
        if (d >= P->dim || g[d] == 0) g[d] = 1;
        P->mapg[d] = (d < P->dim) ? g[d] : 1;
        P->mapn[d] = (d < P->dim) ? P->n[d] / P->mapg[d] : 1;
    }
    std::vector<int> map((size_t)P->mapn[0] * P->mapn[1] * P->mapn[2], -1);
    for (int b = 0; b < P->nboxes; ++b) {
        const BoxDev &B = P->boxes[b];
        int l0 = (B.lo[0] - P->dlo[0]) / P->mapg[0], h0 = (B.hi[0] - P->dlo[0]) / P->mapg[0];
        int l1 = (B.lo[1] - P->dlo[1]) / P->mapg[1], h1 = (B.hi[1] - P->dlo[1]) / P->mapg[1];
        int l2 = (P->dim == 3) ? (B.lo[2] - P->dlo[2]) / P->mapg[2] : 0;
        int h2 = (P->dim == 3) ? (B.hi[2] - P->dlo[2]) / P->mapg[2] : 0;
        for (int k = l2; k <= h2; ++k)
            for (int j = l1; j <= h1; ++j)
                for (int i = l0; i <= h0; ++i) {
                    int &m = map[i + (size_t)P->mapn[0] * (j + (size_t)P->mapn[1] * k)];
                    if (m >= 0) { delete P; return MAC_ERR_LAYOUT; }
                    m = b;
                }
    }
    // extents for launches
    for (const BoxDev &B : P->boxes) {
        int e[3];
        for (int d = 0; d < 3; ++d) e[d] = (d < P->dim) ? B.hi[d] - B.lo[d] + 1 : 1;
        P->maxcells = std::max(P->maxcells, e[0] * e[1] * e[2]);
        P->maxcols = std::max(P->maxcols, e[0] * e[1]);
        P->maxez = std::max(P->maxez, e[2]);
        for (int d = 0; d < P->dim; ++d) {
            int f[3] = {e[0], e[1], e[2]};
            f[d] += 1;
            P->maxfaces = std::max(P->maxfaces, f[0] * f[1] * f[2]);
            int gg[3];
            for (int q = 0; q < 3; ++q) gg[q] = (q < P->dim) ? f[q] + 2 * P->ng_u : 1;
            // ghost faces only: the grown face box minus the valid faces
            P->maxghost = std::max(P->maxghost, gg[0] * gg[1] * gg[2] - f[0] * f[1] * f[2]);
            for (int dir = 0; dir < P->dim; ++dir) {
                long plane = 1;
                for (int q = 0; q < P->dim; ++q)
                    if (q != dir) plane *= gg[q];
                P->maxplane = std::max(P->maxplane, plane);
            }
        }
    }
    P->nr = P->n[P->dim - 1];
    int rc = MAC_OK;
    auto fail = [&](int r) { mac_plan_destroy(P); return r; };
    if (hipGetDevice(&P->device) != hipSuccess) return fail(MAC_ERR_HIP);
    if (hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking) != hipSuccess) return fail(MAC_ERR_HIP);
    for (int e = 0; e < 8; ++e)
        if (hipEventCreateWithFlags(&P->ev[e], hipEventDisableSystemFence) != hipSuccess) return fail(MAC_ERR_HIP);
    if (hipMalloc(&P->boxes_dev, sizeof(BoxDev) * P->nboxes) != hipSuccess) return fail(MAC_ERR_HIP);
    if (hipMemcpy(P->boxes_dev, P->boxes.data(), sizeof(BoxDev) * P->nboxes, hipMemcpyHostToDevice) !=
        hipSuccess)
        return fail(MAC_ERR_HIP);
    if (hipMalloc(&P->boxmap_dev, sizeof(int) * map.size()) != hipSuccess) return fail(MAC_ERR_HIP);
    if (hipMemcpy(P->boxmap_dev, map.data(), sizeof(int) * map.size(), hipMemcpyHostToDevice) != hipSuccess)
        return fail(MAC_ERR_HIP);
    if (hipMalloc(&P->mean_dev, sizeof(double)) != hipSuccess) return fail(MAC_ERR_HIP);
    if (hipMalloc(&P->nrm_dev, sizeof(unsigned long long) * 4) != hipSuccess) return fail(MAC_ERR_HIP);
    if (hipHostMalloc(&P->nrm_host, sizeof(unsigned long long) * 4, 0) != hipSuccess) return fail(MAC_ERR_HIP);
    if (hipMemset(P->nrm_dev, 0, sizeof(unsigned long long) * 4) != hipSuccess) return fail(MAC_ERR_HIP);
    if (hipHostMalloc(&P->nrm_map, 64, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
        hipHostGetDevicePointer((void **)&P->nrm_map_dev, P->nrm_map, 0) != hipSuccess) {
        P->nrm_map_ok = false;   // norms then come back through a plain copy
        (void)hipGetLastError();
    } else {
        memset(P->nrm_map, 0, 64);
    }
    if (hipMalloc(&P->err_dev, sizeof(int)) != hipSuccess) return fail(MAC_ERR_HIP);
    if (hipMalloc(&P->tab_dev, sizeof(double *) * (size_t)P->nboxes * (2 * P->dim + 3)) != hipSuccess)
        return fail(MAC_ERR_HIP);
    P->tab_host.assign((size_t)P->nboxes * (2 * P->dim + 3), nullptr);
    if (hipMalloc(&P->b0_dev, sizeof(double) * (P->nr + 1)) != hipSuccess) return fail(MAC_ERR_HIP);
    if (hipMalloc(&P->b0e_dev, sizeof(double) * (P->nr + 1)) != hipSuccess) return fail(MAC_ERR_HIP);
    if ((rc = build_hier<double>(P, P->H64))) return fail(rc);
    *out = P;
    return MAC_OK;
}

static void free_stage(mac_plan *P)
{
    auto fr = [](std::vector<double *> &v) {
        for (double *p : v)
            if (p) (void)hipFree(p);
        v.clear();
    };
    fr(P->hs.umac); fr(P->hs.phi); fr(P->hs.rho); fr(P->hs.rhs); fr(P->hs.cart);
    P->hs.base_ready = P->hs.rhs_ready = P->hs.cart_ready = false;
    for (int i = 0; i < 2; ++i) {
        if (P->hs.slot[i]) (void)hipHostFree(P->hs.slot[i]);
        if (P->hs.moved[i]) (void)hipEventDestroy(P->hs.moved[i]);
        P->hs.slot[i] = nullptr;
        P->hs.moved[i] = nullptr;
    }
    delete P->hs.pool;
    P->hs.pool = nullptr;
}

extern "C" int mac_plan_profile(mac_plan *P, int enable)
{
    if (!P) return MAC_ERR_INVALID;
    prof_flush(P);
    if (enable && P->prof_ev.empty()) {
        P->prof_ev.resize(2 * PROF_PAIRS);
        P->prof_cls.resize(PROF_PAIRS);
        for (auto &e : P->prof_ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    }
    for (int c = 0; c < MAC_PROF_NCLASS; ++c) { P->prof_cnt[c] = 0; P->prof_ms[c] = 0.0; }
    P->prof_on = enable != 0;
    return MAC_OK;
}

extern "C" int mac_plan_profile_read(mac_plan *P, int cls, long *count, double *total_ms)
{
    if (!P || cls < 0 || cls >= MAC_PROF_NCLASS) return MAC_ERR_INVALID;
    prof_flush(P);
    if (count) *count = P->prof_cnt[cls];
    if (total_ms) *total_ms = P->prof_ms[cls];
    return MAC_OK;
}

extern "C" int mac_plan_destroy(mac_plan *P)
{
    if (!P) return MAC_OK;
    if (P->stream) (void)hipStreamSynchronize(P->stream);
    for (auto &e : P->prof_ev)
        if (e) (void)hipEventDestroy(e);
    free_hier(P->H64);
    free_hier(P->H32);
    for (void *p : P->c16_allocs) (void)hipFree(p);
    free_stage(P);
    if (P->boxes_dev) (void)hipFree(P->boxes_dev);
    if (P->boxmap_dev) (void)hipFree(P->boxmap_dev);
    if (P->part) (void)hipFree(P->part);
    if (P->mean_dev) (void)hipFree(P->mean_dev);
    if (P->nrm_dev) (void)hipFree(P->nrm_dev);
    if (P->nrm_host) (void)hipHostFree(P->nrm_host);
    if (P->nrm_map) (void)hipHostFree(P->nrm_map);
    if (P->err_dev) (void)hipFree(P->err_dev);
    if (P->cg) (void)hipFree(P->cg);
    if (P->tab_dev) (void)hipFree(P->tab_dev);
    if (P->b0_dev) (void)hipFree(P->b0_dev);
    if (P->b0e_dev) (void)hipFree(P->b0e_dev);
    if (P->scratch_cells) (void)hipFree(P->scratch_cells);
    if (P->call_tab) (void)hipFree(P->call_tab);
    for (int e = 0; e < 8; ++e)
        if (P->ev[e]) (void)hipEventDestroy(P->ev[e]);
    if (P->stream) (void)hipStreamDestroy(P->stream);
    delete P;
    return MAC_OK;
}

extern "C" int mac_plan_set_opts(mac_plan *P, const mac_opts *o)
{
    if (!P || !o) return MAC_ERR_INVALID;
    int old_levels = P->opts.max_levels;
    P->opts = *o;
    if (o->max_levels != old_levels) {
        free_hier(P->H64);
        free_hier(P->H32);
        int rc = build_hier<double>(P, P->H64);
        if (rc) return rc;
    }
    return MAC_OK;
}

// --------------------------------------------------------- per-box sizes
static size_t box_elems(const mac_plan *P, int b, int ng, int fd)
{
    const BoxDev &B = P->boxes[b];
    size_t v = 1;
    for (int d = 0; d < P->dim; ++d) v *= (size_t)(B.hi[d] - B.lo[d] + 1 + 2 * ng + (d == fd));
    return v;
}

// ------------------------------------------- host arrays through pinned slots
// A host-pointer call's arrays (pageable memory) move through two pinned
// slots: host threads copy one slot's share of the arrays while the copy
// engine moves the other slot, so the link runs from pinned memory and the
// host copies overlap it.  Small calls (below STAGE_MIN bytes) copy each array
// directly.
static constexpr size_t STAGE_SLOT = 64ull << 20;
static constexpr size_t STAGE_MIN = 16ull << 20;

struct Piece {
    const char *src;
    char *dst;
    size_t n;
};

// copy the pieces with the pool's threads, each an equal share of the bytes
static void pool_copy(CopyPool &pool, const std::vector<Piece> &pc, size_t total)
{
    const int T = pool.size();
    pool.run([&](int t) {
        const size_t lo = total * (size_t)t / (size_t)T, hi = total * (size_t)(t + 1) / (size_t)T;
        size_t at = 0;
        for (const Piece &p : pc) {
            const size_t a = std::max(lo, at), b = std::min(hi, at + p.n);
            if (a < b) memcpy(p.dst + (a - at), p.src + (a - at), b - a);
            at += p.n;
            if (at >= hi) break;
        }
    });
}

// the slots, their events and the copy threads; false where any of them
// cannot be had (the arrays are then copied directly)
static bool stage_pinned(mac_plan *P)
{
    HostStage &S = P->hs;
    if (S.pool) return true;
    for (int i = 0; i < 2; ++i) {
        if (!S.slot[i] && hipHostMalloc((void **)&S.slot[i], STAGE_SLOT, hipHostMallocDefault) != hipSuccess) {
            S.slot[i] = nullptr;
            return false;
        }
        if (!S.moved[i] && hipEventCreateWithFlags(&S.moved[i], hipEventDisableTiming) != hipSuccess) {
            S.moved[i] = nullptr;
            return false;
        }
    }
    const unsigned hw = std::thread::hardware_concurrency();
    try {
        S.pool = new CopyPool((int)std::max(1u, std::min(8u, hw ? hw : 1u)));
    } catch (...) {   // no threads to be had: stay on direct copies
        S.pool = nullptr;
        return false;
    }
    return true;
}

// in: host -> device (src host, dst device); else device -> host
static int move_pieces(mac_plan *P, const std::vector<Piece> &jobs, bool in)
{
    HostStage &S = P->hs;
    const hipMemcpyKind kind = in ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
    size_t total = 0;
    for (const Piece &j : jobs) total += j.n;
    if (total < STAGE_MIN || !stage_pinned(P)) {
        for (const Piece &j : jobs) HIPCHK(hipMemcpyAsync(j.dst, j.src, j.n, kind, P->stream));
        return MAC_OK;
    }
    int rc;
    // a transfer of an earlier call that returned early may still use a slot
    for (int i = 0; i < 2; ++i) HIPCHK(hipEventSynchronize(S.moved[i]));
    // groups of pieces that fill one slot each (an array larger than a slot is split)
    std::vector<std::vector<Piece>> groups(1);
    size_t used = 0;
    for (const Piece &j : jobs) {
        size_t done = 0;
        while (done < j.n) {
            if (used == STAGE_SLOT) {
                groups.emplace_back();
                used = 0;
            }
            const size_t take = std::min(j.n - done, STAGE_SLOT - used);
            groups.back().push_back({j.src + done, j.dst + done, take});
            used += take;
            done += take;
        }
    }
    // the slot side of a group: pieces between the slot and the host arrays
    auto host_side = [&](const std::vector<Piece> &g, int s) {
        std::vector<Piece> hp;
        size_t off = 0;
        for (const Piece &p : g) {
            if (in) hp.push_back({p.src, S.slot[s] + off, p.n});
            else hp.push_back({S.slot[s] + off, p.dst, p.n});
            off += p.n;
        }
        pool_copy(*S.pool, hp, off);
    };
    auto engine = [&](const std::vector<Piece> &g, int s) -> int {
        size_t off = 0;
        for (const Piece &p : g) {
            if (in) HIPCHK(hipMemcpyAsync(p.dst, S.slot[s] + off, p.n, kind, P->stream));
            else HIPCHK(hipMemcpyAsync(S.slot[s] + off, p.src, p.n, kind, P->stream));
            off += p.n;
        }
        HIPCHK(hipEventRecord(S.moved[s], P->stream));
        return MAC_OK;
    };
    const size_t ng = groups.size();
    if (in) {
        // fill slot g & 1 once its previous transfer (group g - 2) has left it
        for (size_t g = 0; g < ng; ++g) {
            const int s = (int)(g & 1);
            if (g >= 2) HIPCHK(hipEventSynchronize(S.moved[s]));
            host_side(groups[g], s);
            if ((rc = engine(groups[g], s))) return rc;
        }
    } else {
        // group g lands in slot g & 1 while the host drains group g - 1
        for (size_t g = 0; g <= ng; ++g) {
            if (g < ng && (rc = engine(groups[g], (int)(g & 1)))) return rc;
            if (g >= 1) {
                const int s = (int)((g - 1) & 1);
                HIPCHK(hipEventSynchronize(S.moved[s]));
                host_side(groups[g - 1], s);
            }
        }
    }
    return MAC_OK;
}

static int stage_in(mac_plan *P, double *const *umac, double *const *phi, const double *const *rho,
                    const double *const *macrhs, const double *const *cart)
{
    HostStage &S = P->hs;
    const int nb = P->nboxes, dim = P->dim;
    if (!S.base_ready) {
        S.umac.assign((size_t)nb * dim, nullptr);
        S.phi.assign(nb, nullptr);
        S.rho.assign(nb, nullptr);
        for (int b = 0; b < nb; ++b) {
            for (int d = 0; d < dim; ++d)
                HIPCHK(hipMalloc(&S.umac[b * dim + d], sizeof(double) * box_elems(P, b, P->ng_u, d)));
            HIPCHK(hipMalloc(&S.phi[b], sizeof(double) * box_elems(P, b, P->ng_phi, -1)));
            HIPCHK(hipMalloc(&S.rho[b], sizeof(double) * box_elems(P, b, P->ng_rho, -1)));
        }
        S.base_ready = true;
    }
    if (macrhs && !S.rhs_ready) {
        S.rhs.assign(nb, nullptr);
        for (int b = 0; b < nb; ++b)
            HIPCHK(hipMalloc(&S.rhs[b], sizeof(double) * box_elems(P, b, P->ng_rhs, -1)));
        S.rhs_ready = true;
    }
    if (cart && !S.cart_ready) {
        S.cart.assign((size_t)nb * dim, nullptr);
        for (int b = 0; b < nb; ++b)
            for (int d = 0; d < dim; ++d)
                HIPCHK(hipMalloc(&S.cart[b * dim + d], sizeof(double) * box_elems(P, b, P->ng_cart, d)));
        S.cart_ready = true;
    }
    std::vector<Piece> jobs;
    auto add = [&](const double *src, double *dst, long n) {
        jobs.push_back({(const char *)src, (char *)dst, sizeof(double) * (size_t)n});
    };
    for (int b = 0; b < nb; ++b) {
        for (int d = 0; d < dim; ++d) add(umac[b * dim + d], S.umac[b * dim + d], box_elems(P, b, P->ng_u, d));
        add(phi[b], S.phi[b], box_elems(P, b, P->ng_phi, -1));
        add(rho[b], S.rho[b], box_elems(P, b, P->ng_rho, -1));
        if (macrhs) add(macrhs[b], S.rhs[b], box_elems(P, b, P->ng_rhs, -1));
        if (cart)
            for (int d = 0; d < dim; ++d) add(cart[b * dim + d], S.cart[b * dim + d], box_elems(P, b, P->ng_cart, d));
    }
    return move_pieces(P, jobs, true);
}

static int stage_out(mac_plan *P, double *const *umac, double *const *phi)
{
    HostStage &S = P->hs;
    std::vector<Piece> jobs;
    auto add = [&](const double *src, double *dst, long n) {
        jobs.push_back({(const char *)src, (char *)dst, sizeof(double) * (size_t)n});
    };
    for (int b = 0; b < P->nboxes; ++b) {
        for (int d = 0; d < P->dim; ++d) add(S.umac[b * P->dim + d], umac[b * P->dim + d], box_elems(P, b, P->ng_u, d));
        add(S.phi[b], phi[b], box_elems(P, b, P->ng_phi, -1));
    }
    int rc = move_pieces(P, jobs, false);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(P->stream));
    return MAC_OK;
}

// copy the level-0 solution back into the caller's per-box phi
__global__ void k_phi_out(const BoxDev *boxes, double *const *phi, int ng, int dim, const double *src,
                          int dlo0, int dlo1, int dlo2, int nx, int ny)
{
    const int b = blockIdx.y;
    const BoxDev B = boxes[b];
    const int ex = B.hi[0] - B.lo[0] + 1, ey = B.hi[1] - B.lo[1] + 1;
    const int ez = (dim == 3) ? B.hi[2] - B.lo[2] + 1 : 1;
    const long q = (long)blockIdx.x * 256 + threadIdx.x;
    if (q >= (long)ex * ey * ez) return;
    int li = (int)(q % ex), lj = (int)((q / ex) % ey), lk = (int)(q / ((long)ex * ey));
    int i = B.lo[0] + li, j = B.lo[1] + lj, k = (dim == 3) ? B.lo[2] + lk : 0;
    long c = (i - dlo0) + (long)nx * ((j - dlo1) + (long)ny * ((dim == 3) ? k - dlo2 : 0));
    long ax = ex + 2 * ng, ay = ey + 2 * ng;
    long o = (li + ng) + ax * ((lj + ng) + ay * ((dim == 3) ? lk + ng : 0));
    phi[b][o] = src[c];
}

static float ev_ms(mac_plan *P, int a, int b)
{
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, P->ev[a], P->ev[b]) != hipSuccess) return 0.f;
    return ms;
}

// --------------------------------------------------------- mac_project
// ------------------------------------------ no-allocation projection scratch
// mac_project_noalloc keeps the per-call device buffers of a projection (the
// caller pointer table, the partial sums of the solvability mean, the mean
// and the error flag) in this file-scope array instead of plan storage that
// grows on demand.  Each GPU holds its own copy for the life of the process;
// in host code the array's name is not a device address, so the address is
// taken with hipGetSymbolAddress once per device and used everywhere else.
// Calls that overlap on one GPU share the array and race.
static constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)2 << 20;   // 2 MiB
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
static constexpr int SCRATCH_MAX_DEVICES = 64;
static unsigned char *scratch_dev[SCRATCH_MAX_DEVICES];
static std::mutex scratch_mu;

// device dev's copy of the array (set up on the first call on that device,
// with dev current)
static int scratch_base(int dev, unsigned char **out)
{
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return MAC_ERR_INVALID;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) return MAC_ERR_HIP;
        scratch_dev[dev] = (unsigned char *)ptr;
    }
    *out = scratch_dev[dev];
    return MAC_OK;
}

// the device buffers one projection call uses
struct ProjBufs {
    double **tab;   // caller pointer table
    double *part;   // partial sums of the solvability mean (pairs)
    double *mean;
    int *err;
};

// whether the solver storage of the plan's precision is allocated (the FP64
// hierarchy at plan creation; the FP32 one, and the compact level where it
// applies, on the first solve in that precision)
static bool storage_ready(mac_plan *P)
{
    if (!P->H64.built) return false;
    if (P->opts.precision == MAC_PREC_FP64) return true;
    if (!P->H32.built) return false;
    return !(compact_applies(P) && P->c16_allocs.empty());
}

// mac_project; noalloc: the call's buffers come from the scratch array and
// nothing is allocated (device pointers only, solver storage already there)
static int project_impl(mac_plan *P, double *const *umac, double *const *phi, const double *const *rho,
                        const double *const *macrhs, const double *beta0_1d, const double *beta0_1d_edge,
                        const double *const *beta0_cart_edge, mac_stats *st, bool noalloc)
{
    if (!P || !umac || !phi || !rho) return MAC_ERR_INVALID;
    if (beta0_1d && beta0_cart_edge) return MAC_ERR_INVALID;      // macproject.f90:99-101
    if (beta0_1d && !beta0_1d_edge) return MAC_ERR_INVALID;
    if (!beta0_1d && beta0_1d_edge) return MAC_ERR_INVALID;
    if (P->opts.precision < MAC_PREC_FP64 || P->opts.precision > MAC_PREC_MIXED_BF16) return MAC_ERR_INVALID;
    if (noalloc && (!P->opts.device_ptrs || !storage_ready(P))) return MAC_ERR_INVALID;
    if (st) memset(st, 0, sizeof(*st));
    hipStream_t s = P->stream;
    const int nb = P->nboxes, dim = P->dim;
    int rc;
    HIPCHK(hipSetDevice(P->device));
    // a planar beta0 lets the FP32 finest level run in the rho form, whose
    // rho the setup writes: the FP32 storage is allocated before it
    P->planar = beta0_1d != nullptr;
    if (!noalloc && (P->opts.precision == MAC_PREC_MIXED || P->opts.precision == MAC_PREC_MIXED_BF16))
        if ((rc = build_hier<float>(P, P->H32))) return rc;
    // the call's device buffers
    const long np = setup_nparts(dim, P->maxcols, P->maxez, nb);
    ProjBufs PB;
    if (noalloc) {
        unsigned char *base = nullptr;
        if ((rc = scratch_base(P->device, &base))) return rc;
        auto up = [](size_t b) { return (b + 255) / 256 * 256; };
        const size_t tab_b = up(sizeof(double *) * P->tab_host.size()), part_b = up(sizeof(double) * 2 * np);
        if (tab_b + part_b + 512 > NOALLOC_SCRATCH_BYTES) return MAC_ERR_TOO_LARGE;
        PB.tab = (double **)base;
        PB.part = (double *)(base + tab_b);
        PB.mean = (double *)(base + tab_b + part_b);
        PB.err = (int *)(base + tab_b + part_b + 256);
    } else {
        if ((rc = ensure_part(P, np))) return rc;
        PB.tab = P->tab_dev;
        PB.part = P->part;
        PB.mean = P->mean_dev;
        PB.err = P->err_dev;
    }
    HIPCHK(hipEventRecord(P->ev[0], s));
    // resolve device pointers
    std::vector<double *> &tab = P->tab_host;
    double **t_umac = tab.data(), **t_phi = t_umac + nb * dim, **t_rho = t_phi + nb,
           **t_rhs = t_rho + nb, **t_cart = t_rhs + nb;
    if (P->opts.device_ptrs) {
        for (int q = 0; q < nb * dim; ++q) t_umac[q] = umac[q];
        for (int b = 0; b < nb; ++b) {
            t_phi[b] = phi[b];
            t_rho[b] = (double *)rho[b];
            if (macrhs) t_rhs[b] = (double *)macrhs[b];
        }
        if (beta0_cart_edge)
            for (int q = 0; q < nb * dim; ++q) t_cart[q] = (double *)beta0_cart_edge[q];
    } else {
        if ((rc = stage_in(P, umac, phi, rho, macrhs, beta0_cart_edge))) return rc;
        for (int q = 0; q < nb * dim; ++q) t_umac[q] = P->hs.umac[q];
        for (int b = 0; b < nb; ++b) {
            t_phi[b] = P->hs.phi[b];
            t_rho[b] = P->hs.rho[b];
            if (macrhs) t_rhs[b] = P->hs.rhs[b];
        }
        if (beta0_cart_edge)
            for (int q = 0; q < nb * dim; ++q) t_cart[q] = P->hs.cart[q];
    }
    HIPCHK(hipMemcpyAsync(PB.tab, tab.data(), sizeof(double *) * tab.size(), hipMemcpyHostToDevice, s));
    if (beta0_1d) {
        HIPCHK(hipMemcpyAsync(P->b0_dev, beta0_1d, sizeof(double) * P->nr,
                              P->opts.device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(P->b0e_dev, beta0_1d_edge, sizeof(double) * (P->nr + 1),
                              P->opts.device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    }
    HIPCHK(hipEventRecord(P->ev[1], s));
    // ---- setup: beta0 U, rh, coefficients, warm start
    SetupArgs A;
    A.dim = dim; A.nboxes = nb; A.boxes = P->boxes_dev;
    for (int d = 0; d < 3; ++d) { A.dlo[d] = P->dlo[d]; A.n[d] = P->n[d]; A.dx[d] = P->dx[d]; }
    A.bc = P->bc;
    A.ng_u = P->ng_u; A.ng_phi = P->ng_phi; A.ng_rho = P->ng_rho; A.ng_rhs = P->ng_rhs; A.ng_cart = P->ng_cart;
    double **td = PB.tab;
    A.umac = td; A.phi = td + nb * dim; A.rho = td + nb * dim + nb;
    A.macrhs = macrhs ? td + nb * dim + 2 * nb : nullptr;
    A.cart = beta0_cart_edge ? td + nb * dim + 3 * nb : nullptr;
    A.beta0 = beta0_1d ? P->b0_dev : nullptr;
    A.beta0e = beta0_1d ? P->b0e_dev : nullptr;
    A.alt = P->opts.use_alt_energy_fix;
    A.L = P->H64.lev[0];
    A.rho32 = rho_mode(P) ? P->H32.rho0.rho : nullptr;
    A.part = PB.part;
    A.err = PB.err;
    HIPCHK(hipMemsetAsync(PB.err, 0, sizeof(int), s));
    int pq = prof_begin(P, MAC_PROF_SETUP);
    launch_setup(A, P->maxcols, P->maxez, s);
    prof_end(P, pq);
    const long n0 = ncell(P->n);
    // max |rh| (the "Initial rhs" figure of the solver log), after the mean
    // removal of a singular problem (fused with it)
    HIPCHK(hipMemsetAsync(P->nrm_dev, 0, sizeof(unsigned long long), s));
    if (P->singular) {
        launch_mean(PB.part, np, (double)n0, PB.mean, s);
        launch_sub_mean_absmax(P->H64.lev[0].rhs, n0, PB.mean, P->nrm_dev, s);
    } else {
        launch_absmax(P->H64.lev[0].rhs, n0, P->nrm_dev, s);
    }
    HIPCHK(hipEventRecord(P->ev[2], s));
    // ---- solve (the non-finite-input flag comes back with the initial
    // residual norm, so the host does not wait for the setup on its own; the
    // solver leaves nrm_dev[0], max |rh|, alone)
    mac_stats local;
    memset(&local, 0, sizeof local);
    P->setup_err = PB.err;
    int src = solve(P, &local);
    P->setup_err = nullptr;
    if (src != MAC_OK && src != MAC_ERR_NOT_CONVERGED) {
        (void)hipStreamSynchronize(s);
        return src;
    }
    HIPCHK(hipEventRecord(P->ev[3], s));
    // ---- correction, ghost fill, phi out
    CorrectArgs C;
    C.dim = dim; C.nboxes = nb; C.boxes = P->boxes_dev;
    for (int d = 0; d < 3; ++d) {
        C.dlo[d] = P->dlo[d]; C.n[d] = P->n[d]; C.dx[d] = P->dx[d];
        C.mapg[d] = P->mapg[d]; C.mapn[d] = P->mapn[d];
        for (int q = 0; q < 2; ++q) C.phys[d][q] = P->phys[d][q];
    }
    C.bc = P->bc;
    C.ng_u = P->ng_u; C.ng_cart = P->ng_cart;
    C.umac = td; C.beta0 = A.beta0; C.beta0e = A.beta0e; C.cart = A.cart;
    const Lev<double> &L0 = P->H64.lev[0];
    C.phi = L0.phi;
    C.bx = L0.bx; C.by = L0.by; C.bz = L0.bz; C.bxh = L0.bxh; C.byh = L0.byh; C.bzh = L0.bzh;
    C.boxmap = P->boxmap_dev;
    C.maxplane = P->maxplane;
    C.phi_out = (double *const *)(td + nb * dim);   // phi out rides on the x-face correction
    C.ng_phi = P->ng_phi;
    pq = prof_begin(P, MAC_PROF_CORRECT);
    launch_correct(C, P->maxcols, P->maxez, P->maxghost, s);
    prof_end(P, pq);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(P->ev[4], s));
    if (!P->opts.device_ptrs) {
        if ((rc = stage_out(P, umac, phi))) return rc;
    }
    HIPCHK(hipEventRecord(P->ev[5], s));
    HIPCHK(hipMemcpyAsync(P->nrm_host + 2, P->nrm_dev, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    memcpy(&local.rhs_norm, P->nrm_host + 2, sizeof(double));
    if (st) {
        *st = local;
        st->t_h2d_ms = ev_ms(P, 0, 1);
        st->t_setup_ms = ev_ms(P, 1, 2);
        st->t_solve_ms = ev_ms(P, 2, 3);
        st->t_correct_ms = ev_ms(P, 3, 4);
        st->t_d2h_ms = ev_ms(P, 4, 5);
        st->t_total_ms = ev_ms(P, 0, 5);
    }
    return src;
}

extern "C" int mac_project(mac_plan *P, double *const *umac, double *const *phi,
                           const double *const *rho, const double *const *macrhs,
                           const double *beta0_1d, const double *beta0_1d_edge,
                           const double *const *beta0_cart_edge, mac_stats *st)
{
    return project_impl(P, umac, phi, rho, macrhs, beta0_1d, beta0_1d_edge, beta0_cart_edge, st, false);
}

extern "C" int mac_project_noalloc(mac_plan *P, double *const *umac, double *const *phi,
                                   const double *const *rho, const double *const *macrhs,
                                   const double *beta0_1d, const double *beta0_1d_edge,
                                   const double *const *beta0_cart_edge, mac_stats *st)
{
    return project_impl(P, umac, phi, rho, macrhs, beta0_1d, beta0_1d_edge, beta0_cart_edge, st, true);
}

// -------------------------------------------------- whole-domain entries
static int level_weights_and_rhs(mac_plan *P, LevelIO &IO)
{
    const Lev<double> &L0 = P->H64.lev[0];
    long n0 = ncell(P->n);
    long np = (n0 + 255) / 256;
    int rc;
    if ((rc = ensure_part(P, np))) return rc;
    launch_level_setup(P->dim, IO, L0, P->part, P->stream);
    if (P->singular) {
        launch_mean(P->part, np, (double)n0, P->mean_dev, P->stream);
        launch_sub_mean<double>(L0.rhs, n0, P->mean_dev, P->stream);
    }
    return MAC_OK;
}

extern "C" int mac_project_level(mac_plan *P, double *ux, double *uy, double *uz, double *phi,
                                 const double *bx, const double *by, const double *bz,
                                 const double *macrhs, mac_stats *st)
{
    if (!P || !ux || !uy || !phi || !bx || !by) return MAC_ERR_INVALID;
    if (P->dim == 3 && (!uz || !bz)) return MAC_ERR_INVALID;
    if (!P->opts.device_ptrs) return MAC_ERR_INVALID;
    if (st) memset(st, 0, sizeof(*st));
    HIPCHK(hipSetDevice(P->device));
    hipStream_t s = P->stream;
    HIPCHK(hipEventRecord(P->ev[0], s));
    LevelIO IO;
    IO.ux = ux; IO.uy = uy; IO.uz = uz;
    IO.bx = bx; IO.by = by; IO.bz = bz;
    IO.macrhs = macrhs; IO.phi = phi;
    for (int d = 0; d < 3; ++d) IO.dx[d] = P->dx[d];
    IO.bc = P->bc;
    int rc;
    if ((rc = level_weights_and_rhs(P, IO))) return rc;
    HIPCHK(hipEventRecord(P->ev[1], s));
    mac_stats local;
    memset(&local, 0, sizeof local);
    P->planar = false;   // the rho form is for mac_project, whose setup writes rho
    int src = solve(P, &local);
    if (src != MAC_OK && src != MAC_ERR_NOT_CONVERGED) return src;
    HIPCHK(hipEventRecord(P->ev[2], s));
    const Lev<double> &L0 = P->H64.lev[0];
    launch_level_correct(P->dim, IO, L0, s);
    HIPCHK(hipMemcpyAsync(phi, L0.phi, sizeof(double) * ncell(P->n), hipMemcpyDeviceToDevice, s));
    HIPCHK(hipEventRecord(P->ev[3], s));
    HIPCHK(hipStreamSynchronize(s));
    if (st) {
        *st = local;
        st->t_setup_ms = ev_ms(P, 0, 1);
        st->t_solve_ms = ev_ms(P, 1, 2);
        st->t_correct_ms = ev_ms(P, 2, 3);
        st->t_total_ms = ev_ms(P, 0, 3);
    }
    return src;
}

extern "C" int mac_adjoint_level(mac_plan *P, const double *phi, const double *bx, const double *by,
                                 const double *bz, const double *g_phi, const double *g_ux,
                                 const double *g_uy, const double *g_uz, double *lambda, double *out_ux,
                                 double *out_uy, double *out_uz, double *out_bx, double *out_by,
                                 double *out_bz, mac_stats *st)
{
    if (!P || !phi || !bx || !by || !g_phi || !g_ux || !g_uy || !lambda || !out_ux || !out_uy ||
        !out_bx || !out_by)
        return MAC_ERR_INVALID;
    if (P->dim == 3 && (!bz || !g_uz || !out_uz || !out_bz)) return MAC_ERR_INVALID;
    if (!P->opts.device_ptrs) return MAC_ERR_INVALID;
    if (st) memset(st, 0, sizeof(*st));
    HIPCHK(hipSetDevice(P->device));
    hipStream_t s = P->stream;
    const long n0 = ncell(P->n);
    const Lev<double> &L0 = P->H64.lev[0];
    HIPCHK(hipEventRecord(P->ev[0], s));
    // weights from beta (no velocity: rh is replaced below), lambda guess 0
    if (!P->scratch_cells) HIPCHK(hipMalloc(&P->scratch_cells, sizeof(double) * n0));
    launch_fill<double>(P->scratch_cells, n0, 0.0, s);
    LevelIO IO;
    memset(&IO, 0, sizeof IO);
    // zero faces for the divergence: reuse the output buffers, cleared
    long nfx = (long)(P->n[0] + 1) * P->n[1] * P->n[2];
    long nfy = (long)P->n[0] * (P->n[1] + 1) * P->n[2];
    long nfz = (long)P->n[0] * P->n[1] * (P->n[2] + 1);
    launch_fill<double>(out_ux, nfx, 0.0, s);
    launch_fill<double>(out_uy, nfy, 0.0, s);
    if (P->dim == 3) launch_fill<double>(out_uz, nfz, 0.0, s);
    IO.ux = out_ux; IO.uy = out_uy; IO.uz = out_uz;
    IO.bx = bx; IO.by = by; IO.bz = bz;
    IO.macrhs = nullptr;
    IO.phi = P->scratch_cells;
    for (int d = 0; d < 3; ++d) IO.dx[d] = P->dx[d];
    IO.bc = P->bc;
    int rc;
    if ((rc = level_weights_and_rhs(P, IO))) return rc;
    // adjoint right side a = g_phi - G^T(w g_U), mean-free on singular problems
    AdjIO J;
    J.gphi = g_phi; J.gux = g_ux; J.guy = g_uy; J.guz = g_uz;
    J.out_rhs = L0.rhs;
    J.out_ux = out_ux; J.out_uy = out_uy; J.out_uz = out_uz;
    J.out_bx = out_bx; J.out_by = out_by; J.out_bz = out_bz;
    for (int d = 0; d < 3; ++d) J.dx[d] = P->dx[d];
    J.bc = P->bc;
    long np = (n0 + 255) / 256;
    launch_adj_rhs(P->dim, J, L0, P->part, s);
    if (P->singular) {
        launch_mean(P->part, np, (double)n0, P->mean_dev, s);
        launch_sub_mean<double>(L0.rhs, n0, P->mean_dev, s);
    }
    HIPCHK(hipEventRecord(P->ev[1], s));
    mac_stats local;
    memset(&local, 0, sizeof local);
    P->planar = false;   // the rho form is for mac_project, whose setup writes rho
    int src = solve(P, &local);
    if (src != MAC_OK && src != MAC_ERR_NOT_CONVERGED) return src;
    if (P->singular)
        if ((rc = remove_mean(P, L0.phi, n0))) return rc;
    HIPCHK(hipEventRecord(P->ev[2], s));
    HIPCHK(hipMemcpyAsync(lambda, L0.phi, sizeof(double) * n0, hipMemcpyDeviceToDevice, s));
    launch_adj_faces(P->dim, J, L0, phi, L0.phi, s);
    HIPCHK(hipEventRecord(P->ev[3], s));
    HIPCHK(hipStreamSynchronize(s));
    if (st) {
        *st = local;
        st->t_setup_ms = ev_ms(P, 0, 1);
        st->t_solve_ms = ev_ms(P, 1, 2);
        st->t_correct_ms = ev_ms(P, 2, 3);
        st->t_total_ms = ev_ms(P, 0, 3);
    }
    return src;
}

// ------------------------------------------------------------ diagnostics
static int debug_array(mac_plan *P, int level, int what, double **ptr, long *n)
{
    if (!P || level < 0 || level >= P->H64.nlev) return MAC_ERR_INVALID;
    const Lev<double> &L = P->H64.lev[level];
    long nc = (long)L.nx * L.ny * L.nz;
    switch (what) {
    case MAC_ARR_PHI: *ptr = L.phi; *n = nc; break;
    case MAC_ARR_RHS: *ptr = L.rhs; *n = nc; break;
    case MAC_ARR_RES: *ptr = L.res; *n = nc; break;
    case MAC_ARR_BX: *ptr = L.bx; *n = nc; break;
    case MAC_ARR_BY: *ptr = L.by; *n = nc; break;
    case MAC_ARR_BZ: *ptr = L.bz; *n = nc; break;
    case MAC_ARR_BXH: *ptr = L.bxh; *n = (long)L.ny * L.nz; break;
    case MAC_ARR_BYH: *ptr = L.byh; *n = (long)L.nx * L.nz; break;
    case MAC_ARR_BZH: *ptr = L.bzh; *n = (long)L.nx * L.ny; break;
    default: return MAC_ERR_INVALID;
    }
    if (!*ptr) return MAC_ERR_INVALID;
    return MAC_OK;
}

extern "C" int mac_debug_level_dims(mac_plan *P, int level, int dims[3])
{
    if (!P || !dims || level < 0 || level >= P->H64.nlev) return MAC_ERR_INVALID;
    dims[0] = P->H64.lev[level].nx;
    dims[1] = P->H64.lev[level].ny;
    dims[2] = P->H64.lev[level].nz;
    return MAC_OK;
}

// the FP32-hierarchy (MAC_ARR_F32) or compact level-0 (MAC_ARR_BF16) copy of
// an array, widened to FP64 on the host
static int debug_get_narrow(mac_plan *P, int level, int what, double *out)
{
    const int base = what & 0xf;
    if (!P->H32.built || level < 0 || level >= P->H32.nlev) return MAC_ERR_INVALID;
    const Lev<float> &F = P->H32.lev[level];
    const long nc = (long)F.nx * F.ny * F.nz;
    const void *src = nullptr;
    long n = nc;
    if (what & MAC_ARR_F32) {
        const float *a[9] = {F.phi, F.rhs, F.bx, F.by, F.bz, F.bxh, F.byh, F.bzh, F.res};
        if (base > 8) return MAC_ERR_INVALID;
        src = a[base];
    } else {
        const Compact16 &K = P->c16;
        if (level != 0 || P->c16_allocs.empty()) return MAC_ERR_INVALID;
        const unsigned short *a[8] = {nullptr, K.rhs, K.bx, K.by, K.bz, K.bxh, K.byh, K.bzh};
        if (base > 7) return MAC_ERR_INVALID;
        src = a[base];
    }
    if (base == MAC_ARR_BXH) n = (long)F.ny * F.nz;
    if (base == MAC_ARR_BYH) n = (long)F.nx * F.nz;
    if (base == MAC_ARR_BZH) n = (long)F.nx * F.ny;
    if (!src) return MAC_ERR_INVALID;
    const size_t es = (what & MAC_ARR_F32) ? 4 : 2;
    std::vector<unsigned char> h(es * (size_t)n);
    HIPCHK(hipMemcpyAsync(h.data(), src, h.size(), hipMemcpyDeviceToHost, P->stream));
    HIPCHK(hipStreamSynchronize(P->stream));
    for (long q = 0; q < n; ++q) {
        float f;
        if (es == 4) {
            memcpy(&f, &h[4 * q], 4);
        } else {
            unsigned short u;
            memcpy(&u, &h[2 * q], 2);
            const unsigned w = (unsigned)u << 16;
            memcpy(&f, &w, 4);
        }
        out[q] = (double)f;
    }
    return MAC_OK;
}

extern "C" int mac_debug_get(mac_plan *P, int level, int what, double *out)
{
    if (!P || !out) return MAC_ERR_INVALID;
    if (what & (MAC_ARR_F32 | MAC_ARR_BF16)) return debug_get_narrow(P, level, what, out);
    double *p;
    long n;
    int rc = debug_array(P, level, what, &p, &n);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(out, p, sizeof(double) * n, hipMemcpyDeviceToHost, P->stream));
    HIPCHK(hipStreamSynchronize(P->stream));
    return MAC_OK;
}

extern "C" int mac_debug_set(mac_plan *P, int level, int what, const double *in)
{
    double *p;
    long n;
    int rc = debug_array(P, level, what, &p, &n);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(p, in, sizeof(double) * n, hipMemcpyHostToDevice, P->stream));
    HIPCHK(hipStreamSynchronize(P->stream));
    return MAC_OK;
}

extern "C" int mac_debug_op(mac_plan *P, int level, int op, int count)
{
    if (!P || level < 0 || level >= P->H64.nlev || count < 0) return MAC_ERR_INVALID;
    if (!(P->opts.omega >= 0 && P->opts.omega < 2)) return MAC_ERR_INVALID;
    Hier<double> &H = P->H64;
    set_omega(H, sweep_omega(P));
    const Lev<double> &L = H.lev[level];
    hipStream_t s = P->stream;
    bool has_next = level + 1 < H.nlev;
    if (op == MAC_OP_GSRB32) {
        int rc;
        if ((rc = build_hier<float>(P, P->H32))) return rc;
        set_omega(P->H32, sweep_omega(P));
        const Lev<float> &F = P->H32.lev[level];
        const long n = (long)L.nx * L.ny * L.nz;
        launch_weights_to_f32(P->dim, L, F, s);
        launch_convert<double, float>(L.phi, F.phi, n, s);
        launch_convert<double, float>(L.rhs, F.rhs, n, s);
        const Compact16 *K = nullptr;
        if (level == 0 && (rc = compact_level0(P, &K, false))) return rc;
        if (K) launch_to_bf16(F.rhs, K->rhs, n, s);
        int cur = 0;
        for (int c = 0; c < count; ++c) {
            if (use_fused(P)) {
                if (!launch_gsrb_fused<float>(F, cur ? F.phi2 : F.phi, cur ? F.phi : F.phi2, SWEEP_PLAIN, s, K))
                    return MAC_ERR_INVALID;
                cur ^= 1;
            } else {
                launch_gsrb<float>(P->dim, F, 0, s);
                launch_gsrb<float>(P->dim, F, 1, s);
            }
        }
        launch_convert<float, double>(cur ? F.phi2 : F.phi, L.phi, n, s);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s));
        return MAC_OK;
    }
    if (op == MAC_OP_GSRB32_RHO) {
        if (level != 0 || !P->H32.built || !P->H32.rho0.rho || !rho_level_ok(P->dim, P->H32.lev[0]))
            return MAC_ERR_INVALID;
        set_omega(P->H32, sweep_omega(P));
        const Lev<float> &F = P->H32.lev[0];
        const long n = (long)L.nx * L.ny * L.nz;
        launch_convert<double, float>(L.phi, F.phi, n, s);
        launch_convert<double, float>(L.rhs, F.rhs, n, s);
        int cur = 0;
        for (int c = 0; c < count; ++c) {
            if (!launch_gsrb_rho(F, P->H32.rho0, cur ? F.phi2 : F.phi, cur ? F.phi : F.phi2, SWEEP_PLAIN, s))
                return MAC_ERR_INVALID;
            cur ^= 1;
        }
        launch_convert<float, double>(cur ? F.phi2 : F.phi, L.phi, n, s);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s));
        return MAC_OK;
    }
    if (op == MAC_OP_GSRB && use_fused(P)) {   // sweeps alternate phi <-> phi2
        for (int c = 0; c < count; ++c)
            launch_gsrb_fused<double>(L, (c & 1) ? L.phi2 : L.phi, (c & 1) ? L.phi : L.phi2, SWEEP_PLAIN, s);
        if (count & 1)
            HIPCHK(hipMemcpyAsync(L.phi, L.phi2, sizeof(double) * (size_t)L.nx * L.ny * L.nz,
                                  hipMemcpyDeviceToDevice, s));
        count = 0;
    }
    for (int c = 0; c < count; ++c) {
        switch (op) {
        case MAC_OP_GSRB:
            if (use_fused(P)) {
                launch_gsrb_fused<double>(L, L.phi, L.phi2, SWEEP_PLAIN, s);
                HIPCHK(hipMemcpyAsync(L.phi, L.phi2, sizeof(double) * (size_t)L.nx * L.ny * L.nz,
                                      hipMemcpyDeviceToDevice, s));
            } else {
                launch_gsrb<double>(P->dim, L, 0, s);
                launch_gsrb<double>(P->dim, L, 1, s);
            }
            break;
        case MAC_OP_RESIDUAL: launch_residual<double>(P->dim, L, L.res, nullptr, s); break;
        case MAC_OP_RESTRICT:
            if (!has_next) return MAC_ERR_INVALID;
            launch_restrict<double>(P->dim, L.res, L.nx, L.ny, H.lev[level + 1], s);
            break;
        case MAC_OP_PROLONG:
            if (!has_next) return MAC_ERR_INVALID;
            launch_prolong<double>(P->dim, L, H.lev[level + 1], P->bc, s);
            break;
        case MAC_OP_COARSEN:
            if (!has_next) return MAC_ERR_INVALID;
            launch_coarsen<double>(P->dim, L, H.lev[level + 1], s);
            break;
        case MAC_OP_VCYCLE:
            if (level != 0) return MAC_ERR_INVALID;
            vcycle<double>(P, H);
            break;
        case MAC_OP_RESID_RESTRICT:
            if (!has_next) return MAC_ERR_INVALID;
            launch_resid_restrict<double>(P->dim, L, L.phi, H.lev[level + 1], s);
            break;
        case MAC_OP_PROLONG_SWEEP:
            if (!has_next) return MAC_ERR_INVALID;
            if (use_fused(P)) {
                int cur = 0;
                prolong_cur<double>(P, H, level, cur);
                fused_sweeps<double>(P, H, level, 1, SWEEP_PLAIN, cur);
                HIPCHK(hipMemcpyAsync(L.phi, L.phi2, sizeof(double) * (size_t)L.nx * L.ny * L.nz,
                                      hipMemcpyDeviceToDevice, s));
            } else {
                launch_prolong<double>(P->dim, L, H.lev[level + 1], P->bc, s);
                launch_gsrb<double>(P->dim, L, 0, s);
                launch_gsrb<double>(P->dim, L, 1, s);
            }
            break;
        default: return MAC_ERR_INVALID;
        }
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    return MAC_OK;
}

// ------------------------------------------ producers of the path's inputs
namespace {

// device mirrors of a call's caller arrays (host pointers when
// opts.device_ptrs == 0) and device pointer tables; freed with the object
struct CallStage {
    mac_plan *P;
    std::vector<void *> mem;
    size_t tab_used = 0;   // pointer slots of P->call_tab handed out
    explicit CallStage(mac_plan *p) : P(p) {}
    // room for `slots` device pointers in the plan's table buffer (grown
    // before any slot is handed out; calls on one plan run in stream order)
    int reserve(size_t slots)
    {
        if (slots * sizeof(double *) <= P->call_tab_cap) return MAC_OK;
        HIPCHK(hipStreamSynchronize(P->stream));
        if (P->call_tab) (void)hipFree(P->call_tab);
        P->call_tab = nullptr;
        P->call_tab_cap = 0;
        HIPCHK(hipMalloc(&P->call_tab, slots * sizeof(double *)));
        P->call_tab_cap = slots * sizeof(double *);
        return MAC_OK;
    }
    ~CallStage()
    {
        if (P->stream) (void)hipStreamSynchronize(P->stream);
        for (void *q : mem) (void)hipFree(q);
    }
    // device view of n doubles at h (copied in when in)
    int view(const double *h, size_t n, bool in, double **d)
    {
        if (!h) { *d = nullptr; return MAC_OK; }
        if (P->opts.device_ptrs) { *d = const_cast<double *>(h); return MAC_OK; }
        void *q = nullptr;
        HIPCHK(hipMalloc(&q, sizeof(double) * std::max<size_t>(n, 1)));
        mem.push_back(q);
        if (in) HIPCHK(hipMemcpyAsync(q, h, sizeof(double) * n, hipMemcpyHostToDevice, P->stream));
        *d = (double *)q;
        return MAC_OK;
    }
    int back(double *h, const double *d, size_t n)
    {
        if (!h || P->opts.device_ptrs) return MAC_OK;
        HIPCHK(hipMemcpyAsync(h, d, sizeof(double) * n, hipMemcpyDeviceToHost, P->stream));
        return MAC_OK;
    }
    // the per-box views back into the caller's host arrays (through the pinned slots)
    int back_all(double *const *h, const std::vector<double *> &views, const std::vector<size_t> &n)
    {
        if (!h || P->opts.device_ptrs) return MAC_OK;
        std::vector<Piece> pc;
        for (size_t a = 0; a < n.size(); ++a)
            if (h[a] && views[a]) pc.push_back({(const char *)views[a], (char *)h[a], sizeof(double) * n[a]});
        return move_pieces(P, pc, false);
    }
    // a device table of device views of per-box arrays (host arrays copied in
    // through the pinned slots when in)
    int table(const double *const *h, const std::vector<size_t> &n, bool in, std::vector<double *> &views,
              double ***tab)
    {
        views.assign(n.size(), nullptr);
        if (!h) { *tab = nullptr; return MAC_OK; }
        int rc;
        if (P->opts.device_ptrs) {
            for (size_t a = 0; a < n.size(); ++a) views[a] = const_cast<double *>(h[a]);
        } else {
            // one allocation for the table's arrays, each at a 256-byte boundary
            auto part = [](size_t cnt) { return (sizeof(double) * std::max<size_t>(cnt, 1) + 255) & ~(size_t)255; };
            size_t total = 0;
            for (size_t a = 0; a < n.size(); ++a) total += part(n[a]);
            void *q = nullptr;
            HIPCHK(hipMalloc(&q, std::max<size_t>(total, 256)));
            mem.push_back(q);
            size_t off = 0;
            for (size_t a = 0; a < n.size(); ++a) {
                if (h[a]) views[a] = (double *)((char *)q + off);
                off += part(n[a]);
            }
        }
        if (in && !P->opts.device_ptrs) {
            std::vector<Piece> pc;
            for (size_t a = 0; a < n.size(); ++a)
                if (h[a]) pc.push_back({(const char *)h[a], (char *)views[a], sizeof(double) * n[a]});
            if ((rc = move_pieces(P, pc, true))) return rc;
        }
        if ((tab_used + n.size()) * sizeof(double *) > P->call_tab_cap) return MAC_ERR_INVALID;
        double **q = (double **)P->call_tab + tab_used;
        tab_used += n.size();
        HIPCHK(hipMemcpyAsync(q, views.data(), sizeof(double *) * n.size(), hipMemcpyHostToDevice, P->stream));
        *tab = q;
        return MAC_OK;
    }
};

SphrGeom sphr_geom(const mac_plan *P, const mac_sphr *s)
{
    SphrGeom G;
    G.dr = s->dr;
    G.nr = s->nr;
    for (int d = 0; d < 3; ++d) {
        G.center[d] = s->center[d];
        G.prob_lo[d] = s->prob_lo[d];
        G.dx[d] = P->dx[d];
    }
    G.interp = s->interp;
    return G;
}

bool sphr_ok(const mac_sphr *s)
{
    return s && s->dr > 0.0 && s->nr >= 3 && s->interp >= 1 && s->interp <= 3;
}

// every cell centre's bin index int(radius / dr) is below nr: the kernels'
// expressions evaluated at the eight corner cells, where the radius (monotone
// in each coordinate's distance from the centre) is largest
bool sphr_covers(const mac_plan *P, const mac_sphr *s)
{
    for (int c = 0; c < 8; ++c) {
        double x[3];
        for (int d = 0; d < 3; ++d) {
            const int i = (c >> d) & 1 ? P->n[d] - 1 : 0;
            x[d] = s->prob_lo[d] + ((double)i + 0.5) * P->dx[d] - s->center[d];
        }
        const double radius = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
        if (!(radius / s->dr < (double)s->nr) || (int)(radius / s->dr) > s->nr - 1) return false;
    }
    return true;
}

}  // namespace

extern "C" int mac_cell_to_edge(mac_plan *P, const double *s0_cell, int nr, double *s0_edge)
{
    if (!P || !s0_cell || !s0_edge || nr < 2) return MAC_ERR_INVALID;
    HIPCHK(hipSetDevice(P->device));
    int rc;
    CallStage S(P);
    double *c, *e;
    if ((rc = S.view(s0_cell, nr, true, &c)) || (rc = S.view(s0_edge, nr + 1, false, &e))) return rc;
    launch_cell_to_edge(c, nr, e, P->stream);
    HIPCHK(hipGetLastError());
    if ((rc = S.back(s0_edge, e, nr + 1))) return rc;
    HIPCHK(hipStreamSynchronize(P->stream));
    return MAC_OK;
}

extern "C" int mac_make_macrhs(mac_plan *P, double *const *macrhs, const mac_macrhs_in *in, const mac_sphr *sph)
{
    if (!P || !macrhs || !in || !in->S_cc || !in->delta_gamma1_term || !in->Sbar || !in->beta0)
        return MAC_ERR_INVALID;
    if (in->ng_S < 0 || in->ng_dg < 0 || in->ng_dp < 0 || in->ng_dchi < 0) return MAC_ERR_LAYOUT;
    if (sph && (P->dim != 3 || !sphr_ok(sph) || !sphr_covers(P, sph))) return MAC_ERR_INVALID;
    const bool dpdt_on = sph ? in->dpdt_factor >= 0.0 : in->dpdt_factor > 0.0;
    if (dpdt_on && (!in->delta_p_term || !in->delta_chi || !in->gamma1bar || !in->p0 || (sph && !in->rho0)))
        return MAC_ERR_INVALID;
    if (dpdt_on && in->dt == 0.0) return MAC_ERR_INVALID;
    HIPCHK(hipSetDevice(P->device));
    const int nb = P->nboxes, dim = P->dim;
    const int nr = sph ? sph->nr : P->n[dim - 1];
    int rc;
    CallStage S(P);
    std::vector<size_t> n_rhs(nb), n_S(nb), n_dg(nb), n_dp(nb), n_dchi(nb);
    for (int b = 0; b < nb; ++b) {
        n_rhs[b] = box_elems(P, b, P->ng_rhs, -1);
        n_S[b] = box_elems(P, b, in->ng_S, -1);
        n_dg[b] = box_elems(P, b, in->ng_dg, -1);
        n_dp[b] = box_elems(P, b, in->ng_dp, -1);
        n_dchi[b] = box_elems(P, b, in->ng_dchi, -1);
    }
    std::vector<double *> v_rhs, v_S, v_dg, v_dp, v_dchi;
    double **t_rhs, **t_S, **t_dg, **t_dp = nullptr, **t_dchi = nullptr;
    if ((rc = S.reserve(5 * (size_t)nb))) return rc;
    // macrhs is staged in as well: its ghost cells are left as the caller had them
    if ((rc = S.table(macrhs, n_rhs, true, v_rhs, &t_rhs)) || (rc = S.table(in->S_cc, n_S, true, v_S, &t_S)) ||
        (rc = S.table(in->delta_gamma1_term, n_dg, true, v_dg, &t_dg)))
        return rc;
    if (dpdt_on && ((rc = S.table(in->delta_p_term, n_dp, true, v_dp, &t_dp)) ||
                    (rc = S.table((const double *const *)in->delta_chi, n_dchi, true, v_dchi, &t_dchi))))
        return rc;
    MacrhsArgs A;
    A.dim = dim;
    A.nboxes = nb;
    A.boxes = P->boxes_dev;
    for (int d = 0; d < 3; ++d) A.dlo[d] = P->dlo[d];
    A.rhs = t_rhs; A.ng_rhs = P->ng_rhs;
    A.S = t_S; A.ng_S = in->ng_S;
    A.dg = t_dg; A.ng_dg = in->ng_dg;
    A.dp = t_dp; A.ng_dp = in->ng_dp;
    A.dchi = t_dchi; A.ng_dchi = in->ng_dchi;
    double *d1[5];
    const double *h1[5] = {in->Sbar, in->beta0, dpdt_on ? in->gamma1bar : nullptr, dpdt_on ? in->p0 : nullptr,
                           (dpdt_on && sph) ? in->rho0 : nullptr};
    for (int a = 0; a < 5; ++a)
        if ((rc = S.view(h1[a], nr, true, &d1[a]))) return rc;
    A.Sbar = d1[0]; A.beta0 = d1[1]; A.gamma1bar = d1[2]; A.p0 = d1[3]; A.rho0 = d1[4];
    A.dt = in->dt;
    A.dpdt = in->dpdt_factor;
    A.dpdt_on = dpdt_on;
    A.predictor = in->is_predictor != 0;
    A.cutoff_coord = in->base_cutoff_coord;
    A.cutoff_density = in->base_cutoff_density;
    if (sph) A.geo = sphr_geom(P, sph);
    else memset(&A.geo, 0, sizeof A.geo);
    launch_make_macrhs(A, P->maxcells, sph != nullptr, P->stream);
    HIPCHK(hipGetLastError());
    if ((rc = S.back_all(macrhs, v_rhs, n_rhs))) return rc;
    if (dpdt_on && (rc = S.back_all(in->delta_chi, v_dchi, n_dchi))) return rc;
    HIPCHK(hipStreamSynchronize(P->stream));
    return MAC_OK;
}

extern "C" int mac_make_s0mac(mac_plan *P, const double *s0, const mac_sphr *sph, double *const *s0mac)
{
    if (!P || !s0 || !s0mac || !sphr_ok(sph) || P->dim != 3 || !sphr_covers(P, sph)) return MAC_ERR_INVALID;
    if (P->ng_cart != 1) return MAC_ERR_LAYOUT;   // make_s0mac assumes one ghost layer (:985-987)
    HIPCHK(hipSetDevice(P->device));
    const int nb = P->nboxes;
    int rc;
    CallStage S(P);
    std::vector<size_t> n_f((size_t)nb * 3);
    int maxfaces = 0;
    for (int b = 0; b < nb; ++b)
        for (int d = 0; d < 3; ++d) {
            n_f[b * 3 + d] = box_elems(P, b, 1, d);
            maxfaces = std::max(maxfaces, (int)n_f[b * 3 + d]);
        }
    std::vector<double *> v_f;
    double **t_f;
    double *d_s0;
    if ((rc = S.reserve(3 * (size_t)nb))) return rc;
    if ((rc = S.table(s0mac, n_f, false, v_f, &t_f)) || (rc = S.view(s0, sph->nr, true, &d_s0))) return rc;
    if (!P->scratch_cells) HIPCHK(hipMalloc(&P->scratch_cells, sizeof(double) * (size_t)ncell(P->n)));
    S0macArgs A;
    A.nboxes = nb;
    A.boxes = P->boxes_dev;
    for (int d = 0; d < 3; ++d) {
        A.dlo[d] = P->dlo[d];
        A.n[d] = P->n[d];
    }
    A.s0 = d_s0;
    A.cells = P->scratch_cells;
    A.out = t_f;
    A.geo = sphr_geom(P, sph);
    launch_make_s0mac(A, maxfaces, P->stream);
    HIPCHK(hipGetLastError());
    if ((rc = S.back_all(s0mac, v_f, n_f))) return rc;
    HIPCHK(hipStreamSynchronize(P->stream));
    return MAC_OK;
}

// ------------------------------------------- solver-level entry: mac_solve
extern "C" int mac_ell_bc(int phys, int comp)
{
    // Source/define_bc_tower.f90:329-399
    if (phys == MAC_PERIODIC) return MAC_ELL_PER;
    switch (phys) {
    case MAC_SLIP_WALL:
    case MAC_NO_SLIP_WALL:
    case MAC_SYMMETRY:
        return MAC_ELL_NEU;
    case MAC_INLET:
        return comp == MAC_COMP_SCALAR ? MAC_ELL_DIR : MAC_ELL_NEU;
    case MAC_OUTLET:
        return comp == MAC_COMP_PRESS ? MAC_ELL_DIR : MAC_ELL_NEU;
    default:
        return 0;
    }
}

namespace {

// the call's BCs, tolerances and alpha for the duration of one mac_solve;
// the plan's projection state comes back on every exit
struct HelmScope {
    mac_plan *P;
    EllBC bc;
    bool singular;
    double eps_rel, eps_abs;
    HelmScope(mac_plan *p, const EllBC &call_bc, double er, double ea)
        : P(p), bc(p->bc), singular(p->singular), eps_rel(p->opts.eps_rel), eps_abs(p->opts.eps_abs)
    {
        P->bc = call_bc;
        P->opts.eps_rel = er;
        P->opts.eps_abs = ea;
        P->helm = true;
    }
    ~HelmScope()
    {
        (void)set_alpha<double>(P, P->H64, false);
        (void)set_alpha<float>(P, P->H32, false);
        P->bc = bc;
        P->singular = singular;
        P->opts.eps_rel = eps_rel;
        P->opts.eps_abs = eps_abs;
        P->helm = false;
    }
};

}  // namespace

extern "C" int mac_solve(mac_plan *P, const mac_solve_in *in, mac_stats *st)
{
    if (!P || !in || !in->rh || !in->phi || !in->beta) return MAC_ERR_INVALID;
    if (in->ng_rh < 0 || in->ng_phi < 0 || in->ng_alpha < 0 || in->ng_beta < 0) return MAC_ERR_LAYOUT;
    if (!(in->eps_rel > 0.0) && !(in->eps_abs > 0.0)) return MAC_ERR_INVALID;
    if (P->opts.precision < MAC_PREC_FP64 || P->opts.precision > MAC_PREC_MIXED_BF16) return MAC_ERR_INVALID;
    const int nb = P->nboxes, dim = P->dim;
    EllBC cb = P->bc;
    bool dir = false;
    for (int d = 0; d < dim; ++d)
        for (int side = 0; side < 2; ++side) {
            const int k = in->ell_bc[d][side];
            if (k != MAC_ELL_PER && k != MAC_ELL_DIR && k != MAC_ELL_NEU) return MAC_ERR_INVALID;
            // the hierarchy's periodic wrap comes from the layout
            if ((k == MAC_ELL_PER) != (P->bc.k[d][side] == ELL_PER)) return MAC_ERR_INVALID;
            cb.k[d][side] = k;
            dir |= k == MAC_ELL_DIR;
        }
    if (dir && in->ng_phi < 1) return MAC_ERR_LAYOUT;   // Dirichlet values live in phi's ghosts
    if (st) memset(st, 0, sizeof(*st));
    HIPCHK(hipSetDevice(P->device));
    hipStream_t s = P->stream;
    int rc;
    HIPCHK(hipEventRecord(P->ev[0], s));
    HelmScope scope(P, cb, in->eps_rel, in->eps_abs);
    CallStage S(P);
    std::vector<size_t> n_rh(nb), n_phi(nb), n_a(nb), n_b((size_t)nb * dim);
    for (int b = 0; b < nb; ++b) {
        n_rh[b] = box_elems(P, b, in->ng_rh, -1);
        n_phi[b] = box_elems(P, b, in->ng_phi, -1);
        n_a[b] = box_elems(P, b, in->ng_alpha, -1);
        for (int d = 0; d < dim; ++d) n_b[b * dim + d] = box_elems(P, b, in->ng_beta, d);
    }
    std::vector<double *> v_rh, v_phi, v_a, v_b;
    double **t_rh, **t_phi, **t_a, **t_b;
    if ((rc = S.reserve((size_t)nb * (3 + dim)))) return rc;
    if ((rc = S.table(in->rh, n_rh, true, v_rh, &t_rh)) ||
        (rc = S.table((const double *const *)in->phi, n_phi, true, v_phi, &t_phi)) ||
        (rc = S.table(in->alpha, n_a, true, v_a, &t_a)) || (rc = S.table(in->beta, n_b, true, v_b, &t_b)))
        return rc;
    if ((rc = set_alpha<double>(P, P->H64, in->alpha != nullptr))) return rc;
    HIPCHK(hipEventRecord(P->ev[1], s));
    // ---- setup: weights, alpha, rh' and warm start from the boxes
    SolveArgs A;
    A.dim = dim;
    A.nboxes = nb;
    A.boxes = P->boxes_dev;
    for (int d = 0; d < 3; ++d) { A.dlo[d] = P->dlo[d]; A.n[d] = P->n[d]; }
    A.bc = P->bc;
    A.rh = t_rh; A.ng_rh = in->ng_rh;
    A.phi = t_phi; A.ng_phi = in->ng_phi;
    A.alpha = t_a; A.ng_alpha = in->ng_alpha;
    A.beta = t_b; A.ng_beta = in->ng_beta;
    A.L = P->H64.lev[0];
    const long np = setup_nparts(2, P->maxcells, 1, nb);
    if ((rc = ensure_part(P, np))) return rc;
    A.part = P->part;
    A.err = P->err_dev;
    HIPCHK(hipMemsetAsync(P->err_dev, 0, sizeof(int), s));
    HIPCHK(hipMemsetAsync(P->nrm_dev + 2, 0, sizeof(unsigned long long), s));
    int pq = prof_begin(P, MAC_PROF_SETUP);
    launch_solve_setup(A, P->maxcells, s);
    // max |alpha| (zero everywhere: the problem may be singular)
    if (in->alpha) launch_absmax(P->H64.lev[0].alpha, ncell(P->n), P->nrm_dev + 2, s);
    prof_end(P, pq);
    HIPCHK(hipGetLastError());
    int err_host = 0;
    unsigned long long amax_bits = 0;
    HIPCHK(hipMemcpyAsync(&err_host, P->err_dev, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&amax_bits, P->nrm_dev + 2, sizeof amax_bits, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (err_host) return MAC_ERR_NONFINITE;
    P->singular = !dir && amax_bits == 0;
    const long n0 = ncell(P->n);
    if (P->singular) {
        launch_mean(P->part, np, (double)n0, P->mean_dev, s);
        launch_sub_mean<double>(P->H64.lev[0].rhs, n0, P->mean_dev, s);
    }
    HIPCHK(hipMemsetAsync(P->nrm_dev, 0, sizeof(unsigned long long), s));
    launch_absmax(P->H64.lev[0].rhs, n0, P->nrm_dev, s);
    unsigned long long rhs_bits = 0;
    HIPCHK(hipMemcpyAsync(&rhs_bits, P->nrm_dev, sizeof rhs_bits, hipMemcpyDeviceToHost, s));
    HIPCHK(hipEventRecord(P->ev[2], s));
    HIPCHK(hipStreamSynchronize(s));
    // ---- solve
    mac_stats local;
    memset(&local, 0, sizeof local);
    P->planar = false;   // the rho form is for mac_project, whose setup writes rho
    const int src = solve(P, &local);
    memcpy(&local.rhs_norm, &rhs_bits, sizeof(double));
    if (src != MAC_OK && src != MAC_ERR_NOT_CONVERGED) return src;
    HIPCHK(hipEventRecord(P->ev[3], s));
    // ---- phi out (valid cells)
    hipLaunchKernelGGL(k_phi_out, dim3((P->maxcells + 255) / 256, nb), dim3(256), 0, s,
                       (const BoxDev *)P->boxes_dev, (double *const *)t_phi, in->ng_phi, dim,
                       (const double *)P->H64.lev[0].phi, P->dlo[0], P->dlo[1], P->dlo[2], P->n[0], P->n[1]);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(P->ev[4], s));
    if ((rc = S.back_all(in->phi, v_phi, n_phi))) return rc;
    HIPCHK(hipEventRecord(P->ev[5], s));
    HIPCHK(hipStreamSynchronize(s));
    if (st) {
        *st = local;
        st->t_h2d_ms = ev_ms(P, 0, 1);
        st->t_setup_ms = ev_ms(P, 1, 2);
        st->t_solve_ms = ev_ms(P, 2, 3);
        st->t_correct_ms = ev_ms(P, 3, 4);
        st->t_d2h_ms = ev_ms(P, 4, 5);
        st->t_total_ms = ev_ms(P, 0, 5);
    }
    return src;
}

// ================================================== two-level projection
// macproject with nlevs = 2 (Source/macproject.f90:51-328): both levels'
// beta0 U, the edge restriction of the fine faces onto the coarse ones
// (:357-361), rh per level and its cell restriction (:389-404), beta per
// level (:178-193), the composite solve (:227-229),
// mkumac per level with the composite fluxes, the edge restriction again
// (:829-833), U / beta0, and the ghost faces: the coarse level's from its
// boxes, periodic images and physical rules, the fine level's from the fine
// boxes, the coarse faces beyond the refined region (create_umac_grown,
// :256-275) and the physical rules.
//
// The composite solve (FBoxLib ml_cc_solve is absent from the reference
// tree) is a fast adaptive composite (FAC) iteration on the composite
// residual: each step restricts the fine residual onto the covered coarse
// cells, corrects with a few V-cycles of the coarse level's plan (whole
// coarse domain) added on both levels, then corrects the fine level with a
// few V-cycles of the fine plan (the region: the fine boxes' bounding
// rectangle, its coarse-fine sides Dirichlet, its holes held near zero).  The composite residual alone decides convergence, so
// the answer is the composite operator's whatever the inner cycles do.
struct mac_ml_plan {
    int dim = 0;
    mac_plan *P[2] = {nullptr, nullptr};   // coarse layout; the fine correction solver (one box: the region)
    mac_plan *PF = nullptr;                 // the caller's fine boxes over the region (tables, ghost fill)
    unsigned char *cov = nullptr;           // the region's covered coarse cells (device)
    bool holes = false;                     // the fine boxes leave part of the region uncovered
    MlGeom G;
    double dxc[3] = {1, 1, 1}, dxf[3] = {1, 1, 1};
    int cphys[3][2];
    double *uc[3] = {nullptr, nullptr, nullptr}, *uf[3] = {nullptr, nullptr, nullptr};
    double *bcf[3] = {nullptr, nullptr, nullptr}, *bff[3] = {nullptr, nullptr, nullptr};
    double *rhc = nullptr, *rhf = nullptr, *phc = nullptr, *phf = nullptr;
    double *mrc = nullptr, *mrf = nullptr, *zero = nullptr, *cvec = nullptr;
    void *block = nullptr;
    int maxgrown = 0;
    mac_opts opts;
};

static constexpr int ML_FP64_CYCLES = 1;     // FP64 inner solves: V-cycles per correction, no norms
static constexpr double ML_CF_WEIGHT = 4.0 / 15.0;   // see k_ml_cf_weights (folded: 2 beta -> 8/15 beta)
static constexpr double ML_HOLE_WEIGHT = 8.0 / 15.0; // the same interface weight on an unfolded face

extern "C" int mac_ml_plan_destroy(mac_ml_plan *M)
{
    if (!M) return MAC_OK;
    if (M->P[0] && M->P[0]->stream) (void)hipStreamSynchronize(M->P[0]->stream);
    if (M->block) (void)hipFree(M->block);
    if (M->cov) (void)hipFree(M->cov);
    if (M->PF) mac_plan_destroy(M->PF);
    if (M->P[1]) {
        M->P[1]->stream = nullptr;   // the coarse plan's stream, destroyed with it
        mac_plan_destroy(M->P[1]);
    }
    if (M->P[0]) mac_plan_destroy(M->P[0]);
    delete M;
    return MAC_OK;
}

extern "C" int mac_ml_plan_create(int nlevs, const mac_layout *lays, const mac_opts *opts, mac_ml_plan **out)
{
    if (!out || !lays) return MAC_ERR_INVALID;
    *out = nullptr;
    if (nlevs != 2) return MAC_ERR_INVALID;
    const mac_layout &L0 = lays[0], &L1 = lays[1];
    const int dim = L0.dim;
    if ((dim != 2 && dim != 3) || L1.dim != dim) return MAC_ERR_LAYOUT;
    if (L1.nboxes < 1 || !L1.box_lo || !L1.box_hi) return MAC_ERR_LAYOUT;
    mac_ml_plan *M = new mac_ml_plan();
    M->dim = dim;
    if (opts) M->opts = *opts;
    else mac_default_opts(&M->opts);
    MlGeom &G = M->G;
    memset(&G, 0, sizeof G);
    G.dim = dim;
    int rlo[3] = {0, 0, 0}, rhi[3] = {0, 0, 0};
    for (int d = 0; d < 3; ++d) {
        if (d < dim) {
            // both domains start at 0; the fine one is the coarse one refined by 2
            if (L0.domain_lo[d] != 0 || L1.domain_lo[d] != 0 || L1.domain_hi[d] != 2 * L0.domain_hi[d] + 1 ||
                std::fabs(L1.dx[d] * 2 - L0.dx[d]) > 1e-12 * L0.dx[d]) {
                delete M;
                return MAC_ERR_LAYOUT;
            }
            for (int s = 0; s < 2; ++s)
                if (L0.phys_bc[d][s] != L1.phys_bc[d][s]) { delete M; return MAC_ERR_LAYOUT; }
            G.nc[d] = L0.domain_hi[d] + 1;
            M->dxc[d] = L0.dx[d];
            M->dxf[d] = L1.dx[d];
            G.ihc[d] = 1.0 / L0.dx[d];
            G.ihf[d] = 1.0 / L1.dx[d];
            rlo[d] = L1.box_lo[d];
            rhi[d] = L1.box_hi[d];
            for (int b = 0; b < L1.nboxes; ++b) {
                rlo[d] = std::min(rlo[d], L1.box_lo[3 * b + d]);
                rhi[d] = std::max(rhi[d], L1.box_hi[3 * b + d]);
                // every box covers whole coarse cells (the fine box plan
                // checks that the boxes are disjoint and inside the domain)
                if ((L1.box_lo[3 * b + d] & 1) || !(L1.box_hi[3 * b + d] & 1)) { delete M; return MAC_ERR_LAYOUT; }
            }
            for (int s = 0; s < 2; ++s) {
                M->cphys[d][s] = L0.phys_bc[d][s];
                G.ckind[d][s] = ell_kind(L0.phys_bc[d][s]);
            }
        } else {
            G.nc[d] = 1;
            G.ihc[d] = G.ihf[d] = 0.0;
            M->cphys[d][0] = M->cphys[d][1] = MAC_PERIODIC;
            G.ckind[d][0] = G.ckind[d][1] = ELL_PER;
        }
        G.rlo[d] = rlo[d];
        G.nr[d] = rhi[d] - rlo[d] + 1;
    }
    // the fine plan: the fine boxes over the region; its coarse-fine sides
    // Dirichlet (the inner correction solves), periodic where the region
    // spans a periodic direction, physical where it meets the domain's side
    mac_layout F = L1;
    for (int d = 0; d < dim; ++d) {
        F.domain_lo[d] = rlo[d];
        F.domain_hi[d] = rhi[d];
        const bool per = G.ckind[d][0] == ELL_PER;
        const bool touch[2] = {rlo[d] == 0, rhi[d] == 2 * G.nc[d] - 1};
        for (int s = 0; s < 2; ++s) {
            if (per && touch[0] && touch[1]) {
                G.side[d][s] = ML_SIDE_PER;
                F.phys_bc[d][s] = MAC_PERIODIC;
            } else if (touch[s] && !per) {
                G.side[d][s] = ML_SIDE_PHYS;
                F.phys_bc[d][s] = L1.phys_bc[d][s];
            } else {
                G.side[d][s] = ML_SIDE_CF;
                F.phys_bc[d][s] = MAC_OUTLET;
            }
        }
    }
    int rc;
    mac_opts po = M->opts;
    po.device_ptrs = 1;
    if ((rc = mac_plan_create(&L0, &po, &M->P[0]))) { mac_ml_plan_destroy(M); return rc; }
    // the caller's fine boxes over the region (holes allowed): box tables,
    // owner map and ghost fill of the fine level
    if ((rc = plan_create(&F, &po, &M->PF, true))) { mac_ml_plan_destroy(M); return rc; }
    // the fine correction solver: one box, the region
    {
        mac_layout R = F;
        int one_lo[3] = {rlo[0], rlo[1], rlo[2]}, one_hi[3] = {rhi[0], rhi[1], rhi[2]};
        R.nboxes = 1;
        R.box_lo = one_lo;
        R.box_hi = one_hi;
        if ((rc = mac_plan_create(&R, &po, &M->P[1]))) { mac_ml_plan_destroy(M); return rc; }
    }
    // the covered coarse cells of the region
    {
        const int cn[3] = {G.nr[0] / 2, std::max(G.nr[1] / 2, 1), std::max(G.nr[2] / 2, 1)};
        std::vector<unsigned char> cov((size_t)cn[0] * cn[1] * cn[2], 0);
        long ncov = 0;
        for (int b = 0; b < L1.nboxes; ++b) {
            int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
            for (int d = 0; d < dim; ++d) {
                lo[d] = (L1.box_lo[3 * b + d] - rlo[d]) / 2;
                hi[d] = (L1.box_hi[3 * b + d] - rlo[d]) / 2;
            }
            for (int k = lo[2]; k <= hi[2]; ++k)
                for (int j = lo[1]; j <= hi[1]; ++j)
                    for (int i = lo[0]; i <= hi[0]; ++i) {
                        cov[i + (size_t)cn[0] * (j + (size_t)cn[1] * k)] = 1;
                        ++ncov;
                    }
        }
        M->holes = ncov < (long)cov.size();
        if (hipMalloc(&M->cov, cov.size()) != hipSuccess ||
            hipMemcpy(M->cov, cov.data(), cov.size(), hipMemcpyHostToDevice) != hipSuccess) {
            mac_ml_plan_destroy(M);
            return MAC_ERR_HIP;
        }
        G.cov = M->holes ? M->cov : nullptr;   // no mask: every region cell covered
    }
    // one stream for both levels
    (void)hipStreamSynchronize(M->P[1]->stream);
    (void)hipStreamDestroy(M->P[1]->stream);
    M->P[1]->stream = M->P[0]->stream;
    // whole-level arrays
    const long ncc = ncell(G.nc), nfc = ncell(G.nr);
    long nfc_c[3], nfc_f[3];
    for (int d = 0; d < 3; ++d) {
        int a[3] = {G.nc[0], G.nc[1], G.nc[2]}, b[3] = {G.nr[0], G.nr[1], G.nr[2]};
        a[d] += 1;
        b[d] += 1;
        nfc_c[d] = (d < dim) ? ncell(a) : 0;
        nfc_f[d] = (d < dim) ? ncell(b) : 0;
    }
    auto up = [](long n) { return (size_t)((n + 31) / 32 * 32); };
    size_t tot = 0;
    for (int d = 0; d < dim; ++d) tot += 2 * up(nfc_c[d]) + 2 * up(nfc_f[d]);
    const long zmax = std::max({nfc_c[0], nfc_c[1], nfc_c[2], nfc_f[0], nfc_f[1], nfc_f[2]});
    tot += 3 * up(ncc) + 3 * up(nfc) + up(zmax) + up(ncc + nfc);
    if (hipMalloc(&M->block, tot * sizeof(double)) != hipSuccess) { mac_ml_plan_destroy(M); return MAC_ERR_NOMEM; }
    double *p = (double *)M->block;
    auto take = [&](long n) { double *q = p; p += up(n); return q; };
    for (int d = 0; d < dim; ++d) {
        M->uc[d] = take(nfc_c[d]);
        M->bcf[d] = take(nfc_c[d]);
        M->uf[d] = take(nfc_f[d]);
        M->bff[d] = take(nfc_f[d]);
    }
    M->rhc = take(ncc); M->phc = take(ncc); M->mrc = take(ncc);
    M->rhf = take(nfc); M->phf = take(nfc); M->mrf = take(nfc);
    M->zero = take(zmax);
    M->cvec = take(ncc + nfc);
    // everything zero once: the fine arrays' entries on holes of the region
    // are never written (zero faces and coefficients there)
    if (hipMemset(M->block, 0, tot * sizeof(double)) != hipSuccess) { mac_ml_plan_destroy(M); return MAC_ERR_HIP; }
    // the largest grown face box of a fine box (ghost faces of the fine level)
    for (const BoxDev &B : M->PF->boxes)
        for (int d = 0; d < dim; ++d) {
            long v = 1;
            for (int e = 0; e < dim; ++e) v *= B.hi[e] - B.lo[e] + 1 + 2 * L1.ng_umac + (e == d);
            M->maxgrown = (int)std::max<long>(M->maxgrown, v);
        }
    *out = M;
    return MAC_OK;
}

extern "C" int mac_ml_plan_set_opts(mac_ml_plan *M, const mac_opts *o)
{
    if (!M || !o) return MAC_ERR_INVALID;
    M->opts = *o;
    return MAC_OK;
}

// the plan's solver options for the inner correction cycles (cycle shape,
// sweeps, bottom solver, over-relaxation) from the two-level call's
static void ml_inner_opts(mac_ml_plan *M, mac_plan *P)
{
    P->opts = M->opts;
    P->opts.device_ptrs = 1;
    P->opts.verbose = 0;
    P->planar = false;
}

// one inner correction solve on level 0 of P (right side written, phi
// cleared), without residual norms (no host round trip): FP64, one V-cycle
// of the FP64 hierarchy; the mixed precisions, one FP32 V-cycle from zero on
// the FP32 copy of the right side, widened back into the FP64 phi (the
// composite residual stays FP64, so the answer does not change)
static int ml_inner(mac_ml_plan *M, mac_plan *P, int *cycles)
{
    if (M->opts.precision == MAC_PREC_FP64) {
        for (int c = 0; c < ML_FP64_CYCLES; ++c) vcycle<double>(P, P->H64);
        *cycles += ML_FP64_CYCLES;
        return MAC_OK;
    }
    const long n = ncell(P->n);
    Lev<double> &L = P->H64.lev[0];
    launch_convert<double, float>(L.rhs, P->H32.lev[0].rhs, n, P->stream);
    vcycle<float>(P, P->H32, true);
    launch_convert<float, double>(P->H32.lev[0].phi, L.phi, n, P->stream);
    *cycles += 1;
    return MAC_OK;
}

// the FP32 hierarchy of an inner solver plan from its FP64 level 0 (weights,
// and alpha where the holes set it), once per call
static int ml_inner_f32(mac_plan *P)
{
    int rc;
    if ((rc = build_hier<float>(P, P->H32))) return rc;
    set_omega(P->H32, sweep_omega(P));
    if ((rc = set_alpha<float>(P, P->H32, P->helm && P->H64.alpha_on))) return rc;
    launch_weights_to_f32(P->dim, P->H64.lev[0], P->H32.lev[0], P->stream);
    coarsen_all<float>(P, P->H32);
    return MAC_OK;
}

// the composite residual's max norm (read back; synchronises the stream)
static int ml_norm(mac_ml_plan *M, const MlState &S, bool fine_only, double *out, int *bad)
{
    mac_plan *P0 = M->P[0];
    hipStream_t s = P0->stream;
    HIPCHK(hipMemsetAsync(P0->nrm_dev, 0, sizeof(unsigned long long), s));
    launch_ml_resid(S, P0->H64.lev[0].rhs, M->P[1]->H64.lev[0].rhs, P0->nrm_dev, fine_only, s);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(P0->nrm_host, P0->nrm_dev, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    if (bad) HIPCHK(hipMemcpyAsync(P0->nrm_host + 1, P0->err_dev, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    memcpy(out, P0->nrm_host, sizeof(double));
    if (bad) {
        int e;
        memcpy(&e, P0->nrm_host + 1, sizeof(int));
        *bad = e;
    }
    return MAC_OK;
}

static int ml_project_dev(mac_ml_plan *M, double *const *const *umac, double *const *const *phi,
                          const double *const *const *rho, const double *const *const *macrhs,
                          const double *const *beta0_1d, const double *const *beta0_1d_edge, mac_stats *st)
{
    const bool b0 = beta0_1d != nullptr;
    if (b0 && (!beta0_1d_edge || !beta0_1d[0] || !beta0_1d[1] || !beta0_1d_edge[0] || !beta0_1d_edge[1]))
        return MAC_ERR_INVALID;
    if (!b0 && beta0_1d_edge) return MAC_ERR_INVALID;
    const bool mr = macrhs != nullptr;
    if (mr && (!macrhs[0] || !macrhs[1])) return MAC_ERR_INVALID;
    if (M->opts.precision < MAC_PREC_FP64 || M->opts.precision > MAC_PREC_MIXED_BF16) return MAC_ERR_INVALID;
    if (st) memset(st, 0, sizeof(*st));
    mac_plan *P0 = M->P[0], *P1 = M->P[1];
    const MlGeom &G = M->G;
    const int dim = M->dim;
    const int rd = dim - 1;
    hipStream_t s = P0->stream;
    HIPCHK(hipSetDevice(P0->device));
    HIPCHK(hipEventRecord(P0->ev[0], s));
    HIPCHK(hipMemsetAsync(P0->err_dev, 0, sizeof(int), s));
    // ---- pointer tables and the gather of each level
    const long ncc = ncell(G.nc), nfc = ncell(G.nr);
    for (int l = 0; l < 2; ++l) {
        mac_plan *P = l ? M->PF : M->P[0];
        const int nb = P->nboxes;
        std::vector<double *> &tab = P->tab_host;
        for (int q = 0; q < nb * dim; ++q) tab[q] = umac[l][q];
        for (int b = 0; b < nb; ++b) {
            tab[nb * dim + b] = phi[l][b];
            tab[nb * dim + nb + b] = (double *)rho[l][b];
            tab[nb * dim + 2 * nb + b] = mr ? (double *)macrhs[l][b] : nullptr;
        }
        HIPCHK(hipMemcpyAsync(P->tab_dev, tab.data(), sizeof(double *) * tab.size(), hipMemcpyHostToDevice, s));
        MlGather A;
        A.boxes = P->boxes_dev;
        for (int d = 0; d < 3; ++d) {
            A.org[d] = l ? G.rlo[d] : 0;
            A.n[d] = l ? G.nr[d] : G.nc[d];
        }
        A.ng_u = P->ng_u; A.ng_rho = P->ng_rho; A.ng_rhs = P->ng_rhs; A.ng_phi = P->ng_phi;
        double **td = P->tab_dev;
        A.umac = td;
        A.phi = td + nb * dim;
        A.rho = td + nb * dim + nb;
        A.macrhs = mr ? td + nb * dim + 2 * nb : nullptr;
        A.b0 = b0 ? beta0_1d[l] : nullptr;
        A.b0e = b0 ? beta0_1d_edge[l] : nullptr;
        A.alt = M->opts.use_alt_energy_fix;
        for (int d = 0; d < 3; ++d) {
            A.u[d] = l ? M->uf[d] : M->uc[d];
            A.beta[d] = l ? M->bff[d] : M->bcf[d];
        }
        A.rhs = l ? M->mrf : M->mrc;
        A.phi_out = l ? M->phf : M->phc;
        A.err = P0->err_dev;
        A.cov = l ? G.cov : nullptr;
        launch_ml_gather(dim, A, nb, P->maxcells, s);
    }
    HIPCHK(hipGetLastError());
    // ---- rh: edge restriction, divergence per level, cell restriction
    launch_ml_edge_restrict(G, M->uc, M->uf, s);
    launch_ml_div(dim, G.nc, M->dxc, M->uc, mr ? M->mrc : nullptr, M->rhc, s);
    launch_ml_div(dim, G.nr, M->dxf, M->uf, mr ? M->mrf : nullptr, M->rhf, s);
    launch_ml_cell_restrict(G, M->rhc, M->rhf, s);
    bool singular = true;
    for (int d = 0; d < dim; ++d)
        for (int q = 0; q < 2; ++q)
            if (G.ckind[d][q] == ELL_DIR) singular = false;
    int rc;
    if (singular) {
        // the volume-weighted composite mean (compensated) off both levels
        long nunc = ncc;
        {
            long cov = 1;
            for (int d = 0; d < dim; ++d) cov *= G.nr[d] / 2;
            nunc -= cov;
        }
        const long n = ncc + nfc, np = (n + 255) / 256;
        if ((rc = ensure_part(P0, np))) return rc;
        launch_ml_comp_vec(G, M->rhc, M->rhf, M->cvec, s);
        launch_partial_sum<double>(M->cvec, n, P0->part, s);
        launch_mean(P0->part, np, (double)nunc * (dim == 3 ? 8.0 : 4.0) + (double)nfc, P0->mean_dev, s);
        launch_sub_mean<double>(M->rhc, ncc, P0->mean_dev, s);
        launch_sub_mean<double>(M->rhf, nfc, P0->mean_dev, s);
    }
    // max |rh| over both levels (covered coarse cells hold fine means)
    HIPCHK(hipMemsetAsync(P0->nrm_dev + 2, 0, sizeof(unsigned long long), s));
    launch_absmax(M->rhc, ncc, P0->nrm_dev + 2, s);
    launch_absmax(M->rhf, nfc, P0->nrm_dev + 2, s);
    // ---- the inner solvers' level 0: folded weights from beta (zero U, zero
    // warm start); each step then writes its right side and clears phi
    for (int l = 0; l < 2; ++l) {
        mac_plan *P = M->P[l];
        LevelIO IO;
        IO.ux = IO.uy = IO.uz = M->zero;
        IO.bx = l ? M->bff[0] : M->bcf[0];
        IO.by = l ? M->bff[1] : M->bcf[1];
        IO.bz = (dim == 3) ? (l ? M->bff[2] : M->bcf[2]) : nullptr;
        IO.macrhs = nullptr;
        IO.phi = M->zero;
        for (int d = 0; d < 3; ++d) IO.dx[d] = l ? M->dxf[d] : M->dxc[d];
        IO.bc = P->bc;
        const long n0 = ncell(P->n);
        if ((rc = ensure_part(P, (n0 + 255) / 256))) return rc;
        launch_level_setup(dim, IO, P->H64.lev[0], P->part, s);
        if (l == 1) launch_ml_cf_weights(G, P->H64.lev[0], ML_CF_WEIGHT, s);
        if (l == 1 && M->holes) {
            // holes of the region: pinned near zero through alpha, their
            // faces with the boxes' cells weighted as interface faces
            if ((rc = set_alpha<double>(P, P->H64, true))) return rc;
            launch_ml_holes(G, P->H64.lev[0], ML_HOLE_WEIGHT, s);
            P->helm = true;
        }
        ml_inner_opts(M, P);
        // the weights stay for the whole call: the hierarchy once
        if (M->opts.precision == MAC_PREC_FP64) {
            set_omega(P->H64, sweep_omega(P));
            coarsen_all<double>(P, P->H64);
        } else if ((rc = ml_inner_f32(P))) {
            return rc;
        }
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(P0->ev[1], s));
    // ---- composite solve
    MlState S;
    S.G = G;
    for (int d = 0; d < 3; ++d) {
        S.bc[d] = M->bcf[d];
        S.bf[d] = M->bff[d];
    }
    S.rhc = M->rhc; S.rhf = M->rhf; S.phc = M->phc; S.phf = M->phf;
    const double eps = M->opts.eps_rel;   // the caller's min(eps_mac 10^(nlevs-1), eps_mac_max), :217
    double r0 = 0, r = 0;
    int bad = 0;
    if ((rc = ml_norm(M, S, false, &r0, &bad))) return rc;
    if (bad || !std::isfinite(r0)) return MAC_ERR_NONFINITE;
    r = r0;
    auto done = [&](double x) { return x <= eps * r0 || (M->opts.eps_abs > 0 && x <= M->opts.eps_abs) || x == 0.0; };
    int it = 0, inner = 0;
    rc = MAC_OK;
    while (!done(r) && it < M->opts.max_iter) {
        // coarse correction: the fine residual restricted onto the covered cells
        Lev<double> &C0 = P0->H64.lev[0];
        launch_ml_cell_restrict(G, C0.rhs, P1->H64.lev[0].rhs, s);
        if (P0->singular) {
            if ((rc = remove_mean(P0, C0.rhs, ncc))) return rc;
        }
        HIPCHK(hipMemsetAsync(C0.phi, 0, sizeof(double) * ncc, s));
        if ((rc = ml_inner(M, P0, &inner))) return rc;
        launch_ml_prolong(G, C0.phi, M->phc, M->phf, s);
        // fine correction on the residual left by it
        Lev<double> &F0 = P1->H64.lev[0];
        HIPCHK(hipMemsetAsync(P0->nrm_dev, 0, sizeof(unsigned long long), s));
        launch_ml_resid(S, C0.rhs, F0.rhs, P0->nrm_dev, true, s);
        HIPCHK(hipMemsetAsync(F0.phi, 0, sizeof(double) * nfc, s));
        if ((rc = ml_inner(M, P1, &inner))) return rc;
        launch_ml_axpy(nfc, F0.phi, M->phf, s);
        if ((rc = ml_norm(M, S, false, &r, nullptr))) return rc;
        ++it;
        if (M->opts.verbose) fprintf(stderr, "maestro_mac: composite step %d  res/res0 = %.3e\n", it, r / r0);
        if (!std::isfinite(r)) return MAC_ERR_NONFINITE;
    }
    const bool conv = done(r);
    HIPCHK(hipEventRecord(P0->ev[2], s));
    // ---- phi on covered coarse cells, mkumac, edge restriction, U / beta0
    launch_ml_cell_restrict(G, M->phc, M->phf, s);
    launch_ml_mkumac(S, M->uc, M->uf, s);
    launch_ml_edge_restrict(G, M->uc, M->uf, s);
    if (b0) {
        launch_ml_unscale(dim, G.nc, 0, M->uc, beta0_1d[0], beta0_1d_edge[0], s);
        launch_ml_unscale(dim, G.nr, G.rlo[rd], M->uf, beta0_1d[1], beta0_1d_edge[1], s);
    }
    // ---- the caller's boxes, then their ghost faces
    for (int l = 0; l < 2; ++l) {
        mac_plan *P = l ? M->PF : M->P[0];
        const int nb = P->nboxes;
        MlScatter A;
        A.boxes = P->boxes_dev;
        for (int d = 0; d < 3; ++d) {
            A.org[d] = l ? G.rlo[d] : 0;
            A.n[d] = l ? G.nr[d] : G.nc[d];
            A.u[d] = l ? M->uf[d] : M->uc[d];
        }
        A.ng_u = P->ng_u; A.ng_phi = P->ng_phi;
        A.umac = P->tab_dev;
        A.phi = P->tab_dev + nb * dim;
        A.phi_in = l ? M->phf : M->phc;
        launch_ml_scatter(dim, A, nb, std::max(P->maxfaces, P->maxcells), s);
        CorrectArgs C;
        memset(&C, 0, sizeof C);
        C.dim = dim; C.nboxes = nb; C.boxes = P->boxes_dev;
        for (int d = 0; d < 3; ++d) {
            C.dlo[d] = P->dlo[d]; C.n[d] = P->n[d]; C.dx[d] = P->dx[d];
            C.mapg[d] = P->mapg[d]; C.mapn[d] = P->mapn[d];
            for (int q = 0; q < 2; ++q)
                C.phys[d][q] = (l && G.side[d][q] == ML_SIDE_CF) ? MAC_INTERIOR : P->phys[d][q];
        }
        C.bc = P->bc;
        C.ng_u = P->ng_u;
        C.umac = P->tab_dev;
        C.boxmap = P->boxmap_dev;
        C.maxplane = P->maxplane;
        launch_fill_ghosts(C, P->maxghost, true, false, s);
        if (l == 1 && P->ng_u > 0) {
            MlGrown Gr;
            Gr.G = G;
            Gr.boxes = P->boxes_dev;
            Gr.ng_u = P->ng_u;
            Gr.umac = P->tab_dev;
            for (int d = 0; d < 3; ++d) Gr.uc[d] = M->uc[d];
            launch_ml_grown(dim, Gr, nb, M->maxgrown, s);
        }
        launch_fill_ghosts(C, P->maxghost, false, true, s);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(P0->ev[3], s));
    HIPCHK(hipMemcpyAsync(P0->nrm_host + 2, P0->nrm_dev + 2, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (st) {
        st->iters = it;
        st->inner_cycles = inner;
        st->converged = conv;
        st->nlevels = 2;
        memcpy(&st->rhs_norm, P0->nrm_host + 2, sizeof(double));
        st->res0 = r0;
        st->res = r;
        st->t_setup_ms = ev_ms(P0, 0, 1);
        st->t_solve_ms = ev_ms(P0, 1, 2);
        st->t_correct_ms = ev_ms(P0, 2, 3);
        st->t_total_ms = ev_ms(P0, 0, 3);
    }
    return conv ? MAC_OK : MAC_ERR_NOT_CONVERGED;
}

// host arrays: each level's per-box arrays (ghosts included) and planar
// profiles copied into one device allocation per call, the device path run
// on them, umac and phi copied back
extern "C" int mac_ml_project(mac_ml_plan *M, double *const *const *umac, double *const *const *phi,
                              const double *const *const *rho, const double *const *const *macrhs,
                              const double *const *beta0_1d, const double *const *beta0_1d_edge, mac_stats *st)
{
    if (!M || !umac || !phi || !rho || !umac[0] || !umac[1] || !phi[0] || !phi[1] || !rho[0] || !rho[1])
        return MAC_ERR_INVALID;
    const bool b0 = beta0_1d != nullptr;
    if (b0 && (!beta0_1d_edge || !beta0_1d[0] || !beta0_1d[1] || !beta0_1d_edge[0] || !beta0_1d_edge[1]))
        return MAC_ERR_INVALID;
    if (!b0 && beta0_1d_edge) return MAC_ERR_INVALID;
    const bool mr = macrhs != nullptr;
    if (mr && (!macrhs[0] || !macrhs[1])) return MAC_ERR_INVALID;
    if (M->opts.device_ptrs) return ml_project_dev(M, umac, phi, rho, macrhs, beta0_1d, beta0_1d_edge, st);
    const int dim = M->dim;
    hipStream_t s = M->P[0]->stream;
    HIPCHK(hipSetDevice(M->P[0]->device));
    // sizes (doubles) of every array, in one list per level
    struct Arr { const double *host; double *dev; size_t n; bool back; };
    std::vector<Arr> arrs;
    std::vector<double *> du[2], dp[2], dr[2], dm[2];
    auto cells = [&](const BoxDev &B, int ng, int fd) {
        size_t v = 1;
        for (int e = 0; e < dim; ++e) v *= (size_t)(B.hi[e] - B.lo[e] + 1 + 2 * ng + (e == fd));
        return v;
    };
    size_t tot = 0;
    for (int l = 0; l < 2; ++l) {
        mac_plan *P = l ? M->PF : M->P[0];
        const int nb = P->nboxes;
        for (int b = 0; b < nb; ++b) {
            const BoxDev &B = P->boxes[b];
            for (int d = 0; d < dim; ++d) arrs.push_back({umac[l][b * dim + d], nullptr, cells(B, P->ng_u, d), true});
            arrs.push_back({phi[l][b], nullptr, cells(B, P->ng_phi, -1), true});
            arrs.push_back({rho[l][b], nullptr, cells(B, P->ng_rho, -1), false});
            if (mr) arrs.push_back({macrhs[l][b], nullptr, cells(B, P->ng_rhs, -1), false});
        }
        if (b0) {
            const int nr = (l ? 2 : 1) * M->G.nc[dim - 1];
            arrs.push_back({beta0_1d[l], nullptr, (size_t)nr, false});
            arrs.push_back({beta0_1d_edge[l], nullptr, (size_t)nr + 1, false});
        }
    }
    for (Arr &a : arrs) {
        if (!a.host) return MAC_ERR_INVALID;
        tot += (a.n + 31) / 32 * 32;
    }
    double *blk = nullptr;
    if (hipMalloc(&blk, tot * sizeof(double)) != hipSuccess) return MAC_ERR_NOMEM;
    size_t off = 0;
    int rc = MAC_OK;
    for (Arr &a : arrs) {
        a.dev = blk + off;
        off += (a.n + 31) / 32 * 32;
        if (hipMemcpyAsync(a.dev, a.host, a.n * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess) rc = MAC_ERR_HIP;
    }
    if (rc == MAC_OK) {
        size_t q = 0;
        const double *db0[2] = {nullptr, nullptr}, *db0e[2] = {nullptr, nullptr};
        for (int l = 0; l < 2; ++l) {
            const int nb = (l ? M->PF : M->P[0])->nboxes;
            for (int b = 0; b < nb; ++b) {
                for (int d = 0; d < dim; ++d) du[l].push_back(arrs[q++].dev);
                dp[l].push_back(arrs[q++].dev);
                dr[l].push_back(arrs[q++].dev);
                if (mr) dm[l].push_back(arrs[q++].dev);
            }
            if (b0) {
                db0[l] = arrs[q++].dev;
                db0e[l] = arrs[q++].dev;
            }
        }
        double *const *u2[2] = {du[0].data(), du[1].data()};
        double *const *p2[2] = {dp[0].data(), dp[1].data()};
        const double *const *r2[2] = {dr[0].data(), dr[1].data()};
        const double *const *m2[2] = {dm[0].data(), dm[1].data()};
        rc = ml_project_dev(M, u2, p2, r2, mr ? m2 : nullptr, b0 ? db0 : nullptr, b0 ? db0e : nullptr, st);
        if (rc == MAC_OK || rc == MAC_ERR_NOT_CONVERGED)
            for (Arr &a : arrs)
                if (a.back && hipMemcpyAsync((void *)a.host, a.dev, a.n * sizeof(double), hipMemcpyDeviceToHost, s) !=
                                  hipSuccess)
                    rc = MAC_ERR_HIP;
    }
    if (hipStreamSynchronize(s) != hipSuccess && rc == MAC_OK) rc = MAC_ERR_HIP;
    (void)hipFree(blk);
    return rc;
}
