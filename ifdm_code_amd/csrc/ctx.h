// ctx.h -- the library context (opaque ifdm_ctx of include/ifdm.h) and the
// error / profiling helpers shared by the translation units of libifdm.so.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ifdm.h"

struct ifdm_ctx {
    ifdm_config cfg{};
    int n = 0;
    long n3 = 0;
    // psi, phi and the adjoint fields hold x-plane after x-plane, each plane
    // N*N contiguous cells, planes ps cells apart (ps = N*N + pad); np = N*ps
    long ps = 0, np = 0;
    // half spectrum: rows of hs cells, N rows per x-plane, planes sps apart
    int hs = 0;
    long sps = 0;
    hipStream_t stream = nullptr;
    void *psi = nullptr;       // N^3 complex (plane stride ps)
    void *phi = nullptr;       // N^3 real (plane stride ps)
    void *shat = nullptr;      // N planes of N*HS complex (plane stride sps)
    void *tw = nullptr;        // N complex twiddles (storage precision)
    void *tw64 = nullptr;      // N complex FP64 twiddles (the same buffer in FP64 mode)
    double *k2 = nullptr;      // N doubles
    double *part_rho = nullptr;
    double *part_phi = nullptr;
    double *red = nullptr;     // 3 doubles on device
    double *red_stage = nullptr;  // 3 x IFDM_RED_BLOCKS doubles
    void *dtab = nullptr;      // N complex: per-step drift table
    double *red_host = nullptr;  // 3 doubles, pinned
    double *pk_scratch = nullptr;  // device address of this device's static P(k) scratch
    int nz_blocks = 0;
    int rho_parts = 0;   // density partials written by the last source pass
    // physics
    double m_over_hbar = 0;    // (mfdm / c^2) / hbar
    double g_over_hbar = 0;
    double ms = 0;             // mfdm_solar
    double ke_coef = 0;
    double green = 0;
    double fftnorm = 0;
    // mixed precision: psi is stored as 2^-psi_scale psi (complex64)
    bool mixed = false;
    int psi_scale = 0;
    double psc2 = 1.0;         // 2^(2 psi_scale)
    double rho_mean = 0.0;     // m_s <|psi|^2>, FP64, subtracted in the mixed-mode source
    // second half-kick of the last ifdm_step, not yet applied
    bool pending = false;
    double pend_aphi = 0.0, pend_arho = 0.0;
    // reductions after the last ifdm_step (valid while its half-kick is pending)
    ifdm_stats run_stats{};
    bool run_stats_valid = false;
    // initial-condition grids (ifdm_ic_begin .. ifdm_ic_finish): the FP64
    // density and the velocity grids, float (ic_f32, the reference-order
    // deposit) or FP64 (the atomic deposit)
    double *ic_rho = nullptr;
    void *ic_vx = nullptr, *ic_vy = nullptr, *ic_vz = nullptr;
    bool ic_f32 = true;
    int ic_mode = 0;                  // "ic_deposit" option: 0 reference order, 1 FP64 atomics
    unsigned *ic_first = nullptr, *ic_last = nullptr;   // per-cell ranges of the ordered deposit
    unsigned long long *ic_skipped = nullptr;
    double ic_box = 0.0, ic_pmass = 0.0;
    long long ic_npart = 0;
    // adjoint: gradient field, the real source s and its Poisson image,
    // and the forward checkpoints (psi every ck_k steps, on the device)
    void *adj_g = nullptr, *adj_s = nullptr, *adj_phs = nullptr;
    void *ck = nullptr;
    int ck_k = 0, ck_count = 0, ck_nsteps = -1;
    std::vector<ifdm_dt> ck_sched;   // the schedule the checkpoints were taken on
    // the on-device step loop of ifdm_simulate_ex (option "device_loop"):
    // its state, per-step records (IFDM_REC doubles each, up to
    // IFDM_REC_CAP steps between host waits), the pieces of the quadrature's
    // bisection, and a pinned copy of state and records
    int dev_loop = 1;
    bool dev_active = false;       // the launches being issued belong to the loop
    void *dstep = nullptr;         // one ifdm::DevStep
    double *drec = nullptr;
    void *dpcs = nullptr;          // 3 x IFDM_QAGS_LIMIT ifdm::Piece
    double *hrec = nullptr;        // pinned: DevStep, then the records
    // strided passes with persistent prefetching work-groups: bits 0, 1 the
    // y, x passes on psi; bits 2, 3, 4 the passes F, G, H on the half spectrum
    int persistent = 0;
    int num_cu = 0;
    // optional live per-pass timing (ifdm_set_profiling)
    bool prof = false;
    unsigned prof_mask = ~0u;                 // passes (bit = pass id) that live profiling brackets
    std::vector<hipEvent_t> ev_pool;          // 2 events per profiled launch
    std::vector<int> ev_pass;                 // pass id of each recorded launch
    int cur_pass = 0;                         // pass id of the launches being issued
    double pass_ms[IFDM_NPASS] = {};
    long pass_count = 0;
    std::string err;
};

static constexpr int IFDM_RED_BLOCKS = 256;
static constexpr int IFDM_REC_CAP = 512;

#define HIPCHK(ctx, expr)                                                                  \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);                \
            return IFDM_E_HIP;                                                             \
        }                                                                                  \
    } while (0)

#define LAUNCHCHK(ctx) HIPCHK(ctx, hipGetLastError())

// Copies nx x-planes between a dense array (planes N*N cells apart) and a
// device field of plane stride c->ps, on the context's stream; elem = bytes
// per cell, to_field selects the direction.
static inline hipError_t plane_copy(const ifdm_ctx *c, void *field, void *dense, size_t elem, int nx, bool to_field,
                                    hipMemcpyKind kind) {
    const size_t pb = elem * (size_t)c->n * c->n, fb = elem * (size_t)c->ps;
    if (fb == pb)
        return to_field ? hipMemcpyAsync(field, dense, pb * nx, kind, c->stream)
                        : hipMemcpyAsync(dense, field, pb * nx, kind, c->stream);
    for (int x = 0; x < nx; ++x) {
        char *f = (char *)field + fb * x, *d = (char *)dense + pb * x;
        const hipError_t e = to_field ? hipMemcpyAsync(f, d, pb, kind, c->stream) : hipMemcpyAsync(d, f, pb, kind, c->stream);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Host-side element loops over whole fields (the precision conversions of
// the host copies: 2^28 elements at 512^3): fn(lo, hi, k) over
// IFDM_HOST_CHUNKS fixed pieces of [0, n), k the piece's index, on up to 16
// threads.  The pieces do not depend on the thread count, so per-piece
// partial results combined in piece order give the same bits on any host.
// Runs serially for small n or when no thread can be started.
static constexpr int IFDM_HOST_CHUNKS = 64;
template <typename F> static void host_chunks(size_t n, F fn) {
    auto piece = [&](int k) {
        fn(n * (size_t)k / IFDM_HOST_CHUNKS, n * (size_t)(k + 1) / IFDM_HOST_CHUNKS, k);
    };
    const unsigned hw = std::thread::hardware_concurrency();
    const int nt = (int)std::min<unsigned>(16u, hw ? hw : 1u);
    std::vector<std::thread> th;
    if (nt > 1 && n >= ((size_t)1 << 20)) {
        try {
            for (int t = 1; t < nt; ++t)
                th.emplace_back([&piece, t, nt] { for (int k = t; k < IFDM_HOST_CHUNKS; k += nt) piece(k); });
        } catch (...) {
            // fewer threads: the pieces of the missing ones run below
        }
    }
    const int started = (int)th.size() + 1;
    for (int k = 0; k < IFDM_HOST_CHUNKS; ++k)
        if (k % nt == 0 || k % nt >= started) piece(k);
    for (auto &x : th) x.join();
}

static constexpr int IFDM_EV_POOL = 4096;

// With live profiling on, brackets the next launch with a pair of events
// from the pool and remembers its pass id.
static inline int prof_begin(ifdm_ctx *c) {
    if (!c->prof || !((c->prof_mask >> c->cur_pass) & 1u)) return 0;
    const size_t k = c->ev_pass.size();
    if (2 * k + 1 >= c->ev_pool.size()) { c->err = "profiling event pool exhausted"; return IFDM_E_STATE; }
    HIPCHK(c, hipEventRecord(c->ev_pool[2 * k], c->stream));
    c->ev_pass.push_back(c->cur_pass);
    return 0;
}
static inline int prof_end(ifdm_ctx *c) {
    if (!c->prof || !((c->prof_mask >> c->cur_pass) & 1u)) return 0;
    HIPCHK(c, hipEventRecord(c->ev_pool[2 * (c->ev_pass.size() - 1) + 1], c->stream));
    return 0;
}


// applies a pending second half-kick (defined in ifdm_capi.hip)
extern "C" __attribute__((visibility("hidden"))) int ifdm_internal_flush(ifdm_ctx *c);

// theta solve for the initial conditions: in-place forward 3-D FFT of the
// FP64 complex field buf, multiply by coef / k^2 / N^3 (k = 0 -> 0),
// inverse FFT (src/gadget_IC.h:222-244).  Defined in ifdm_capi.hip.
extern "C" __attribute__((visibility("hidden"))) int ifdm_internal_theta_solve(ifdm_ctx *c, void *buf64, double coef);
