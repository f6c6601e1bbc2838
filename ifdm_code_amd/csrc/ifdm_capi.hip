// ifdm_capi.hip -- the C ABI of libifdm.so (declared in include/ifdm.h):
// context set-up, device buffers, the launch sequence of the fused passes and
// the host-side step loop.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/ifdm.h"
#include "constants.h"
#include "ctx.h"
#include "kdk_passes.h"
#include "ic_kernels.h"

using namespace ifdm;

static thread_local std::string g_create_err;

// Scratch of ifdm_power_spectrum_noalloc: the binned sums (3 per bin) and
// the normalised output (3 per bin), at most IFDM_PK_MAX_BINS bins.  Each
// GPU holds its own copy for the life of the process; every call on a GPU
// shares it (calls that overlap on one GPU race).  Host code uses the
// device address from hipGetSymbolAddress, never the array's name.
#define IFDM_PK_MAX_BINS 4096
static __device__ __align__(256) double scratch[6 * IFDM_PK_MAX_BINS];
static constexpr int IFDM_MAX_DEVICES = 64;
static double *g_scratch_ptr[IFDM_MAX_DEVICES];
static std::mutex g_scratch_mu;

// the device address of `scratch` on the current device `dev`, resolved
// once per device
static hipError_t scratch_address(int dev, double **out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (dev < 0 || dev >= IFDM_MAX_DEVICES) return hipErrorInvalidDevice;
    if (!g_scratch_ptr[dev]) {
        double *p = nullptr;
        const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

namespace {

template <typename T, int N>
struct Ops {
    static constexpr int E = SchedT<T, N>::E, P = N / E;

    static ZArgs zargs(ifdm_ctx *c, double aphi, double arho) {
        ZArgs z;
        z.aphi = aphi;
        z.arho = arho * c->psc2;
        z.ms = c->ms * c->psc2;
        z.rhoc = c->mixed ? c->rho_mean : c->cfg.rhoc;
        z.ref_range = c->cfg.ref_maxphi_range;
        z.blk0 = 0;
        z.dtab = nullptr;
        z.rev = 0;
        z.ps = c->ps;
        z.sps = c->sps;
        z.drhoc = nullptr;
        return z;
    }
    static SArgs sargs(ifdm_ctx *c, double dKE) {
        SArgs s;
        s.other0 = 0;
        s.dtab = c->dtab;
        s.ke_coef = c->ke_coef;
        s.dKE = dKE;
        s.green = c->green;
        s.fftnorm = c->fftnorm;
        s.kunit2 = std::pow(M_PI / c->cfg.xmax, 2);
        s.rev = 0;
        s.aphi = 0.0;
        s.arho = 0.0;
        s.phi = c->phi;
        s.dk = nullptr;
        return s;
    }

    // z pass over x-planes [x0, x0 + nx) (all planes when nx == 0)
    template <int MODE>
    static int zpass(ifdm_ctx *c, ZArgs za, double *part, int x0 = 0, int nx = 0) {
        constexpr int LZ = z_lines<T, N>();
        if (nx == 0) nx = N;
        const long l0 = (long)x0 * N, nl = (long)nx * N;
        za.blk0 = (int)(l0 / LZ);
        const int nb = (int)((nl + LZ - 1) / LZ);
        int rc;
        if ((rc = prof_begin(c))) return rc;
        hipLaunchKernelGGL((k_zpass<T, N, MODE>), dim3(nb), dim3(LZ * P), 0, c->stream,
                           (cx<T> *)c->psi, (T *)c->phi, (cx<T> *)c->shat, (const cx<T> *)c->tw, (const cx<double> *)c->tw64, za, part);
        LAUNCHCHK(c);
        return prof_end(c);
    }
    // pass I over x-planes [x0, x0 + nx): the half-length c2r for N >= 32
    static int c2r(ifdm_ctx *c, int x0 = 0, int nx = 0, int rev = 0) {
        ZArgs za = zargs(c, 0, 0);
        za.rev = rev;
        if constexpr (N < 32) {
            return zpass<Z_C2R>(c, za, c->part_phi, x0, nx);
        } else {
            constexpr int LZ = c2r_lines<T, N>();
            if (nx == 0) nx = N;
            const long l0 = (long)x0 * N, nl = (long)nx * N;
            za.blk0 = (int)(l0 / LZ);
            const int nb = (int)((nl + LZ - 1) / LZ);
            int rc;
            if ((rc = prof_begin(c))) return rc;
            hipLaunchKernelGGL((k_c2r_half<T, N>), dim3(nb), dim3(z_lines<T, N>() * P), 0, c->stream,
                               (const cx<T> *)c->shat, (T *)c->phi, (const cx<T> *)c->tw, za, c->part_phi);
            LAUNCHCHK(c);
            return prof_end(c);
        }
    }

    // strided pass; for AXIS 1 the fixed index (x) runs over [o0, o0 + no)
    // WIDE: 256-byte row tiles in FP64 (the x pass on psi; the half spectrum
    // keeps 128-byte tiles, whose last tile holds a single valid column)
    template <int AXIS, int MODE, bool WIDE = false>
    // rl: row stride, ps: plane stride, kv: valid columns per row
    static int spass(ifdm_ctx *c, void *arr, int rl, long ps, int kv, SArgs sa, int o0 = 0, int no = 0) {
        if (no == 0) no = N;
        if ((MODE == S_DRIFT1 || MODE == S_KICK_DRIFT1) && kv != N) {
            c->err = "drift pass: every column of psi is transformed (kv = N)";
            return IFDM_E_ARG;
        }
        sa.other0 = o0;
        int rc;
        if ((rc = prof_begin(c))) return rc;
        bool done = false;
        // (persistent work-groups spill registers at N = 1024: up to N = 512;
        // complex floats at 1024 need 190-238 VGPRs, or spill 260-412 bytes per
        // lane under the 1024-thread cap of 128)
        if constexpr (MODE != S_KICK_DRIFT1 && N <= 512) {
          constexpr bool PSI = MODE == S_DRIFT1;
          constexpr int bit = PSI ? (AXIS == 1 ? 1 : 2) : (AXIS == 0 ? 8 : (MODE == S_INV ? 16 : 4));
          if (c->persistent & bit) {
            done = true;
            constexpr int TC = s_cols<T, N, false, MODE, AXIS>();
            const int ntz = (kv + TC - 1) / TC;
            static int occ = -1;   // resident work-groups per CU for this instantiation
            if (occ < 0) {
                int o = 0;
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, k_spass_pp<T, N, AXIS, MODE>, TC * P, 0) !=
                        hipSuccess || o < 1)
                    o = 1;
                occ = o;
            }
            const int ntiles = ntz * no;
            const int nb = std::min(ntiles, occ * c->num_cu);
            hipLaunchKernelGGL((k_spass_pp<T, N, AXIS, MODE>), dim3(nb), dim3(TC * P), 0, c->stream,
                               (cx<T> *)arr, rl, ps, kv, (const cx<T> *)c->tw, (const cx<double> *)c->tw64, sa, ntz, ntiles);
          }
        }
        if (!done) {
            constexpr int TC = s_cols<T, N, WIDE, MODE, AXIS>();
            int ntz = (kv + TC - 1) / TC;
            if constexpr (s_narrow<T, N, WIDE, AXIS, MODE>()) {
                // a tile count rounded up to the XCD group (the extra tiles
                // exit at once), and a grid of a multiple of 8 groups
                constexpr int G = xcd_group<MODE>();
                ntz = (ntz + G - 1) / G * G;
                if (no % 8) { c->err = "narrow strided pass: line count must be a multiple of 8"; return IFDM_E_ARG; }
            }
            hipLaunchKernelGGL((k_spass<T, N, AXIS, MODE, WIDE>), dim3(ntz, no), dim3(TC * P), 0, c->stream,
                               (cx<T> *)arr, rl, ps, kv, (const cx<T> *)c->tw, (const cx<double> *)c->tw64, sa);
        }
        LAUNCHCHK(c);
        return prof_end(c);
    }

    // The drift factor of the step separates by axis (it is exp(-i c k^2
    // dKE) with k^2 = k_x^2 + k_y^2 + k_z^2), so each axis runs its own
    // transform - factor - inverse round trip and psi returns to real space
    // after three passes instead of five; the kick rides on the first of
    // them and the Poisson source with its z transform on the last.  One
    // pass of the step by index:
    //   0 T  the per-axis drift table E[i] = exp(-i c k_i^2 dKE) / N
    //   1 A  y: kick (aphi, arho; both 0: none), forward, E[k_y]/N, inverse
    //   2 B  x: forward, E[k_x]/N, inverse
    //   3 C  z: forward, E[k_z]/N, inverse; source s = m_s|psi|^2 - rho_c, its
    //        real-to-complex transform into the half spectrum; max and sum of
    //        |psi|^2
    //   4 F, 5 G, 6 H, 7 I: the Poisson solve on the half spectrum -> phi
    //   8    final reduction of the partials
    // F and I walk their planes in descending order (the pass before each
    // of them ascends), starting on what is still in the memory-side cache.
    static int pass(ifdm_ctx *c, int idx, double aphi, double arho, double dKE) {
        const SArgs sa = sargs(c, dKE);
        SArgs sr = sa;
        sr.rev = 1;
        c->cur_pass = idx;
        switch (idx) {
            case 0: {
                int rc;
                if ((rc = prof_begin(c))) return rc;
                hipLaunchKernelGGL((k_drift_table<T, N>), dim3(1), dim3(256), 0, c->stream, (cx<T> *)c->dtab,
                                   (const double *)c->k2, c->ke_coef * dKE);
                LAUNCHCHK(c);
                return prof_end(c);
            }
            case 1: {
                // 128-byte tiles here (two work-groups per CU in FP64 overlap the
                // kick's sincos with loads: measured faster than 256-byte tiles)
                SArgs sk = sa;
                if (c->dev_active) {
                    // the on-device loop: the coefficients the step preparation wrote
                    sk.dk = &((const DevStep *)c->dstep)->kphi;
                    return spass<1, S_KICK_DRIFT1>(c, c->psi, N, c->ps, N, sk);
                }
                if (aphi == 0.0 && arho == 0.0) return spass<1, S_DRIFT1>(c, c->psi, N, c->ps, N, sa);
                sk.aphi = aphi;
                sk.arho = arho * c->psc2;
                return spass<1, S_KICK_DRIFT1>(c, c->psi, N, c->ps, N, sk);
            }
            // 128-byte tiles with non-temporal loads (0.935 -> 0.91 ms in FP64
            // against 256-byte tiles, which keep plain loads)
            case 2: return spass<0, S_DRIFT1>(c, c->psi, N, c->ps, N, sa);
            case 3: {
                ZArgs za = zargs(c, 0, 0);
                za.dtab = c->dtab;
                if (c->dev_active && c->mixed) za.drhoc = &((const DevStep *)c->dstep)->rho_mean;
                c->rho_parts = c->nz_blocks;
                return zpass<Z_DRIFT_SRC>(c, za, c->part_rho);
            }
            case 4: return spass<1, S_FWD>(c, c->shat, c->hs, c->sps, N / 2 + 1, sr);
            case 5: return spass<0, S_GREEN>(c, c->shat, c->hs, c->sps, N / 2 + 1, sa);
            case 6: return spass<1, S_INV>(c, c->shat, c->hs, c->sps, N / 2 + 1, sa);
            case 7: return c2r(c, 0, 0, 1);
            case 8: {
                int rc;
                if ((rc = prof_begin(c))) return rc;
                hipLaunchKernelGGL(k_reduce1, dim3(IFDM_RED_BLOCKS), dim3(256), 0, c->stream, c->part_rho,
                                   c->rho_parts, c->part_phi, c->nz_blocks, c->red_stage);
                LAUNCHCHK(c);
                // (the on-device loop's next step preparation does the final stage)
                if (!c->dev_active) {
                    hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(256), 0, c->stream, c->red_stage, IFDM_RED_BLOCKS,
                                       c->red);
                    LAUNCHCHK(c);
                }
                return prof_end(c);
            }
            default: return IFDM_E_ARG;
        }
    }

    // phi_dst = G(m_s r) for a real field r (k = 0 -> 0): the adjoint's Poisson solve
    static int poisson_real(ifdm_ctx *c, void *src, void *dst) {
        void *save = c->phi;
        c->phi = src;
        int rc = zpass<Z_RSRC>(c, zargs(c, 0, 0), nullptr);
        c->phi = dst;
        for (int i = 4; i <= 7 && !rc; ++i) rc = pass(c, i, 0, 0, 0);
        c->phi = save;
        return rc;
    }

    // reverse of one half-kick (aphi, arho) on (psi, adj_g) with the current phi
    static int adj_kick(ifdm_ctx *c, double aphi, double arho) {
        // (elementwise over the padded extent: pads hold zeros)
        const unsigned nb = (unsigned)std::min<long>((c->np + 255) / 256, 65536);
        hipLaunchKernelGGL((k_adj_src<T>), dim3(nb), dim3(256), 0, c->stream, (const cx<T> *)c->psi,
                           (const cx<T> *)c->adj_g, (T *)c->adj_s, c->np);
        LAUNCHCHK(c);
        int rc;
        if ((rc = poisson_real(c, c->adj_s, c->adj_phs))) return rc;
        hipLaunchKernelGGL((k_adj_kick<T>), dim3(nb), dim3(256), 0, c->stream, (cx<T> *)c->psi, (cx<T> *)c->adj_g,
                           (const T *)c->phi, (const T *)c->adj_phs, (const T *)c->adj_s, c->np, aphi,
                           arho * c->psc2);
        LAUNCHCHK(c);
        return 0;
    }

    // forward 3-D real transform of m_s|psi|^2 - rho_c into shat, binned
    // into dev_out (3 nbins doubles), cleared by a fill kernel when
    // fill_clear (no memset node for callers that capture the stream)
    static int power_spectrum(ifdm_ctx *c, int nbins, double *dev_out, bool fill_clear) {
        const SArgs sa = sargs(c, 0.0);
        int rc;
        c->rho_parts = c->nz_blocks;
        if ((rc = zpass<Z_SRC>(c, zargs(c, 0, 0), c->part_rho))) return rc;
        if ((rc = spass<1, S_FWD>(c, c->shat, c->hs, c->sps, N / 2 + 1, sa))) return rc;
        if ((rc = spass<0, S_FWD>(c, c->shat, c->hs, c->sps, N / 2 + 1, sa))) return rc;
        if (fill_clear) {
            hipLaunchKernelGGL(k_fill0, dim3((3 * nbins + 255) / 256), dim3(256), 0, c->stream, dev_out, 3 * nbins);
            LAUNCHCHK(c);
        } else {
            HIPCHK(c, hipMemsetAsync(dev_out, 0, sizeof(double) * 3 * nbins, c->stream));
        }
        const long nh = (long)N * N * (N / 2 + 1);
        const unsigned nb = (unsigned)std::min<long>((nh + 255) / 256, 2048);
        hipLaunchKernelGGL((k_pk_bin<T, N>), dim3(nb), dim3(256), sizeof(double) * 3 * nbins, c->stream,
                           (const cx<T> *)c->shat, c->hs, c->sps, nbins, dev_out);
        LAUNCHCHK(c);
        c->cur_pass = 8;
        return pass(c, 8, 0, 0, 0);   // reductions (the norm gives the mean density)
    }

    static int poisson(ifdm_ctx *c) {
        int rc;
        c->rho_parts = c->nz_blocks;
        if ((rc = zpass<Z_SRC>(c, zargs(c, 0, 0), c->part_rho))) return rc;
        for (int i = 4; i <= 8; ++i)
            if ((rc = pass(c, i, 0, 0, 0))) return rc;
        return 0;
    }

    // passes T, A..I + reduction: kick (aphi, arho), drift dKE, Poisson.
    // With profiling on, each pass is bracketed by events on the stream.
    static int kick_drift_poisson(ifdm_ctx *c, double aphi, double arho, double dKE) {
        int rc;
        for (int i = 0; i < IFDM_NPASS; ++i)
            if ((rc = pass(c, i, aphi, arho, dKE))) return rc;
        return 0;
    }

    // IC phase: forward FFT, coef/k^2/N^3 (k = 0 -> 0), inverse FFT of c->psi
    static int theta_solve(ifdm_ctx *c, double coef) {
        SArgs sa = sargs(c, 0.0);
        sa.green = coef;
        const ZArgs z0 = zargs(c, 0, 0);
        int rc;
        if ((rc = zpass<Z_FWD>(c, z0, nullptr))) return rc;
        if ((rc = spass<1, S_FWD, true>(c, c->psi, N, c->ps, N, sa))) return rc;
        if ((rc = spass<0, S_GREEN, true>(c, c->psi, N, c->ps, N, sa))) return rc;
        if ((rc = spass<1, S_INV, true>(c, c->psi, N, c->ps, N, sa))) return rc;
        return zpass<Z_INV>(c, z0, nullptr);
    }

    // the drift alone: table, then the y, x and z round trips
    static int drift(ifdm_ctx *c, double dKE) {
        int rc;
        if ((rc = pass(c, 0, 0.0, 0.0, dKE))) return rc;
        if ((rc = pass(c, 1, 0.0, 0.0, dKE))) return rc;
        if ((rc = pass(c, 2, 0.0, 0.0, dKE))) return rc;
        ZArgs za = zargs(c, 0, 0);
        za.dtab = c->dtab;
        return zpass<Z_DRIFT>(c, za, nullptr);
    }

    static int fft3d(ifdm_ctx *c, int sign) {
        SArgs sa = sargs(c, 0.0);
        int rc;
        if (sign < 0) {
            if ((rc = zpass<Z_FWD>(c, zargs(c, 0, 0), nullptr))) return rc;
            if ((rc = spass<1, S_FWD, true>(c, c->psi, N, c->ps, N, sa))) return rc;
            return spass<0, S_FWD, true>(c, c->psi, N, c->ps, N, sa);
        }
        if ((rc = zpass<Z_INV>(c, zargs(c, 0, 0), nullptr))) return rc;
        if ((rc = spass<1, S_INV, true>(c, c->psi, N, c->ps, N, sa))) return rc;
        return spass<0, S_INV, true>(c, c->psi, N, c->ps, N, sa);
    }

    // one step of the on-device loop: the step preparation (as pass 0, in
    // place of the drift table), the passes A..I with the device
    // coefficients, and the first reduction stage; rec: the step's record
    static int dev_step(ifdm_ctx *c, const StepIn &in, double *rec) {
        c->cur_pass = 0;
        int rc;
        if ((rc = prof_begin(c))) return rc;
        hipLaunchKernelGGL((k_step_prep<T, N>), dim3(1), dim3(256), 0, c->stream, (const double *)c->red_stage,
                           IFDM_RED_BLOCKS, in, (DevStep *)c->dstep, rec, (cx<T> *)c->dtab, (const double *)c->k2,
                           (Piece *)c->dpcs);
        LAUNCHCHK(c);
        if ((rc = prof_end(c))) return rc;
        c->dev_active = true;
        for (int i = 1; i < IFDM_NPASS && !rc; ++i) rc = pass(c, i, 0.0, 0.0, 0.0);
        c->dev_active = false;
        return rc;
    }

    static int kick(ifdm_ctx *c, double aphi, double arho) {
        long nb = (c->np + 255) / 256;
        if (nb > 65536) nb = 65536;
        hipLaunchKernelGGL((k_kick<T>), dim3((unsigned)nb), dim3(256), 0, c->stream, (cx<T> *)c->psi,
                           (const T *)c->phi, c->np, aphi, arho * c->psc2);
        LAUNCHCHK(c);
        return 0;
    }
};

struct Dispatch {
    int (*power_spectrum)(ifdm_ctx *, int, double *, bool);
    int (*adj_kick)(ifdm_ctx *, double, double);
    int (*theta_solve)(ifdm_ctx *, double);
    int (*pass)(ifdm_ctx *, int, double, double, double);
    int (*poisson)(ifdm_ctx *);
    int (*kdp)(ifdm_ctx *, double, double, double);
    int (*drift)(ifdm_ctx *, double);
    int (*fft3d)(ifdm_ctx *, int);
    int (*kick)(ifdm_ctx *, double, double);
    int (*dev_step)(ifdm_ctx *, const StepIn &, double *);
};

template <typename T, int N>
Dispatch make_dispatch() {
    return Dispatch{&Ops<T, N>::power_spectrum, &Ops<T, N>::adj_kick, &Ops<T, N>::theta_solve, &Ops<T, N>::pass, &Ops<T, N>::poisson, &Ops<T, N>::kick_drift_poisson, &Ops<T, N>::drift, &Ops<T, N>::fft3d,
                    &Ops<T, N>::kick, &Ops<T, N>::dev_step};
}

bool get_dispatch(int n, int precision, Dispatch *d) {
    if (precision == IFDM_MIXED) {
        switch (n) {
            case 8: *d = make_dispatch<float, 8>(); return true;
            case 16: *d = make_dispatch<float, 16>(); return true;
            case 32: *d = make_dispatch<float, 32>(); return true;
            case 64: *d = make_dispatch<float, 64>(); return true;
            case 128: *d = make_dispatch<float, 128>(); return true;
            case 256: *d = make_dispatch<float, 256>(); return true;
            case 512: *d = make_dispatch<float, 512>(); return true;
            case 1024: *d = make_dispatch<float, 1024>(); return true;
            default: return false;
        }
    }
    if (precision != IFDM_FP64) return false;
    switch (n) {
        case 8: *d = make_dispatch<double, 8>(); return true;
        case 16: *d = make_dispatch<double, 16>(); return true;
        case 32: *d = make_dispatch<double, 32>(); return true;
        case 64: *d = make_dispatch<double, 64>(); return true;
        case 128: *d = make_dispatch<double, 128>(); return true;
        case 256: *d = make_dispatch<double, 256>(); return true;
        case 512: *d = make_dispatch<double, 512>(); return true;
        case 1024: *d = make_dispatch<double, 1024>(); return true;
        default: return false;
    }
}

// adds the elapsed time of every launch recorded since the last call to
// its pass and counts one step
// Any change of psi from outside the checkpointed run leaves the checkpoints
// without the state they end in: they no longer serve a reverse sweep.
inline void drop_checkpoints(ifdm_ctx *c) { c->ck_nsteps = -1; }

int accumulate_profile(ifdm_ctx *c, int steps = 1) {
    if (!c->prof) return 0;
    const size_t k = c->ev_pass.size();
    if (k == 0) return 0;
    HIPCHK(c, hipEventSynchronize(c->ev_pool[2 * k - 1]));
    for (size_t i = 0; i < k; ++i) {
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev_pool[2 * i], c->ev_pool[2 * i + 1]));
        c->pass_ms[c->ev_pass[i]] += ms;
    }
    c->ev_pass.clear();
    c->pass_count += steps;
    return 0;
}

int read_stats(ifdm_ctx *c, ifdm_stats *st) {
    HIPCHK(c, hipMemcpyAsync(c->red_host, c->red, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const double maxrho = c->red_host[1] * c->psc2, norm = c->red_host[2] * c->psc2;
    if (c->mixed) c->rho_mean = c->ms * norm / (double)c->n3;
    if (st) {
        st->max_phi = c->red_host[0];
        st->max_rho = maxrho;
        st->norm = norm;
    }
    return 0;
}

// k-vector of src/struct3D.h:76-85
double kvec(int i, int n, double xmax) {
    if (i < n / 2) return i * M_PI / xmax;
    return (static_cast<double>(i) - n) * M_PI / xmax;
}

void release(ifdm_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->cfg.device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &e : c->ev_pool)
        if (e) (void)hipEventDestroy(e);
    void *icb[] = {c->ic_rho, c->ic_vx, c->ic_vy, c->ic_vz, c->ic_skipped, c->ic_first, c->ic_last};
    for (void *b : icb)
        if (b) (void)hipFree(b);
    void *bufs[] = {c->psi, c->phi, c->shat, c->tw, c->k2, c->part_rho, c->part_phi, c->red, c->red_stage, c->dtab,
                    c->adj_g, c->adj_s, c->adj_phs, c->ck, c->dstep, c->drec, c->dpcs};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (c->tw64 && c->tw64 != c->tw) (void)hipFree(c->tw64);
    if (c->red_host) (void)hipHostFree(c->red_host);
    if (c->hrec) (void)hipHostFree(c->hrec);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

bool pow2_in_range(int n) { return n >= 8 && n <= 1024 && (n & (n - 1)) == 0; }

// twiddles W_N^k = exp(-2 pi i k / N), k < N, evaluated in long double
void host_twiddles(int n, std::vector<double> &tw, std::vector<float> &twf) {
    tw.resize(2 * (size_t)n);
    twf.resize(2 * (size_t)n);
    for (int k = 0; k < n; ++k) {
        const long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)n;
        tw[2 * k] = (double)cosl(a);
        tw[2 * k + 1] = (double)-sinl(a);
        twf[2 * k] = (float)cosl(a);
        twf[2 * k + 1] = (float)-sinl(a);
    }
}

}  // namespace

extern "C" {

const char *ifdm_create_error(void) { return g_create_err.c_str(); }

const char *ifdm_last_error(const ifdm_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int ifdm_create(const ifdm_config *cfg, ifdm_ctx **out) {
    g_create_err.clear();
    if (!cfg || !out) { g_create_err = "null argument"; return IFDM_E_ARG; }
    *out = nullptr;
    if (!pow2_in_range(cfg->n)) { g_create_err = "n must be a power of two in [8, 1024]"; return IFDM_E_ARG; }
    if (cfg->precision != IFDM_FP64 && cfg->precision != IFDM_MIXED) {
        g_create_err = "precision must be IFDM_FP64 or IFDM_MIXED";
        return IFDM_E_ARG;
    }
    if (!(cfg->xmax > 0) || !(cfg->mfdm > 0)) { g_create_err = "xmax and mfdm must be positive"; return IFDM_E_ARG; }
    if (cfg->plane_pad > 65536) { g_create_err = "plane_pad must be at most 65536"; return IFDM_E_ARG; }
    Dispatch d;
    if (!get_dispatch(cfg->n, cfg->precision, &d)) { g_create_err = "unsupported size"; return IFDM_E_ARG; }

    ifdm_ctx *c = new ifdm_ctx();
    c->cfg = *cfg;
    c->n = cfg->n;
    c->n3 = (long)cfg->n * cfg->n * cfg->n;
    c->hs = hs_of(cfg->n);
    c->mixed = cfg->precision == IFDM_MIXED;
    // x-planes of psi and phi are padded beyond their N*N cells: a
    // power-of-two plane stride sends the rows of an x-line to the same HBM
    // channels (measured at 512^3 with 32 cells: the x pass B 1.09 -> 0.96 ms in FP64,
    // 0.59 -> 0.44 ms mixed; 15.4 -> 10.5 ms FP64 at 1024^3).  Pads round up
    // to 8 cells, so rows stay 8-cell aligned, as the phi lines that pass I
    // stores as complex pairs need.  Built-in: 256 bytes of psi, 16 cells in
    // FP64 and 32 in mixed (a sweep of 16, 32 and 64 cells at 512^3 with
    // the non-temporal accesses: FP64 213.5, 211.7, 208.9 steps/s, pass B
    // 0.86, 0.90, 0.96 ms; mixed within 0.5 % across the three).
    const long dpad = c->mixed ? 32 : 16;
    long pad = cfg->plane_pad > 0 ? cfg->plane_pad : (cfg->plane_pad < 0 ? 0 : dpad);
    pad = (pad + 7) / 8 * 8;
    c->ps = (long)cfg->n * cfg->n + pad;
    c->np = (long)cfg->n * c->ps;
    // (the half spectrum's plane stride N*HS is no power of two: padding it
    // measured neutral)
    c->sps = (long)cfg->n * c->hs;
    // one work-group per tile on every strided pass: persistent work-groups
    // (option "persistent") measured slower or equal at 512^3 (FP64 G 0.459
    // against 0.463 ms, H 0.385 against 0.408)
    c->persistent = 0;
    const int n = c->n;
    auto fail = [&](int code, const std::string &m) {
        g_create_err = m;
        release(c);
        return code;
    };
    hipError_t e = hipSetDevice(cfg->device);
    if (e != hipSuccess) return fail(IFDM_E_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
    if ((e = hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, cfg->device)) != hipSuccess)
        return fail(IFDM_E_HIP, std::string("hipDeviceGetAttribute: ") + hipGetErrorString(e));
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess)
        return fail(IFDM_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    if ((e = scratch_address(cfg->device, &c->pk_scratch)) != hipSuccess)
        return fail(IFDM_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    // element sizes of the stored fields (reductions stay FP64)
    const size_t rs = c->mixed ? sizeof(float) : sizeof(double), cs = 2 * rs, ds = sizeof(double);
    // z-pass block count: lines per block = min(256 / (n / E), n * n)
    const int per = std::min(256 / (n / pts_of(c->mixed, n)), n * n);
    c->nz_blocks = (n * n + per - 1) / per;
    struct { void **p; size_t bytes; } allocs[] = {
        {&c->psi, cs * (size_t)c->np},
        {&c->phi, rs * (size_t)c->np},
        {&c->shat, cs * (size_t)n * c->sps},
        {&c->tw, cs * (size_t)n},
        {(void **)&c->k2, ds * (size_t)n},
        {(void **)&c->part_rho, 2 * ds * (size_t)c->nz_blocks},
        {(void **)&c->part_phi, ds * (size_t)c->nz_blocks},
        {(void **)&c->red, 3 * ds},
        {(void **)&c->red_stage, 3 * ds * IFDM_RED_BLOCKS},
        {&c->dtab, cs * (size_t)n},
        {&c->dstep, sizeof(DevStep)},
        {(void **)&c->drec, ds * IFDM_REC * (size_t)IFDM_REC_CAP},
        {&c->dpcs, sizeof(Piece) * 3 * (size_t)IFDM_QAGS_LIMIT},
    };
    for (auto &a : allocs) {
        if ((e = hipMalloc(a.p, a.bytes)) != hipSuccess)
            return fail(IFDM_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    if ((e = hipHostMalloc((void **)&c->red_host, 3 * ds)) != hipSuccess ||
        (e = hipHostMalloc((void **)&c->hrec, sizeof(DevStep) + ds * IFDM_REC * (size_t)IFDM_REC_CAP)) != hipSuccess)
        return fail(IFDM_E_NOMEM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
    // twiddles W_N^k = exp(-2 pi i k / N) evaluated in long double
    std::vector<double> tw, k2(n);
    std::vector<float> twf;
    host_twiddles(n, tw, twf);
    for (int k = 0; k < n; ++k) k2[k] = std::pow(kvec(k, n, cfg->xmax), 2);
    const void *twsrc = c->mixed ? (const void *)twf.data() : (const void *)tw.data();
    if (!c->mixed) {
        c->tw64 = c->tw;
    } else if ((e = hipMalloc(&c->tw64, 2 * ds * n)) != hipSuccess ||
               (e = hipMemcpy(c->tw64, tw.data(), 2 * ds * n, hipMemcpyHostToDevice)) != hipSuccess) {
        c->tw64 = nullptr;
        return fail(IFDM_E_HIP, std::string("twiddles: ") + hipGetErrorString(e));
    }
    if ((e = hipMemcpy(c->tw, twsrc, cs * n, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(c->k2, k2.data(), ds * n, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemset(c->psi, 0, cs * (size_t)c->np)) != hipSuccess ||
        (e = hipMemset(c->phi, 0, rs * (size_t)c->np)) != hipSuccess ||
        (e = hipMemset(c->shat, 0, cs * (size_t)n * c->sps)) != hipSuccess)
        return fail(IFDM_E_HIP, std::string("init copy: ") + hipGetErrorString(e));
    const double mt = cfg->mfdm / std::pow(c_light, 2);
    c->m_over_hbar = mt / hbar;
    c->g_over_hbar = cfg->g / hbar;
    c->ms = cfg->mfdm * kg2eV / kg2solar;                   // src/CosmoBEC_MPI_stan.cpp:330
    c->ke_coef = 0.5 * hbar / mt;                            // src/struct3D.h:121
    c->green = -4.0 * M_PI * bigG;                           // src/CosmoBEC_MPI_stan.cpp:105
    c->fftnorm = 1. / static_cast<double>(std::pow(n, 3));   // src/struct3D.h:69
    if (c->mixed) {
        // |psi| ~ sqrt(rho_c / m_s) ~ 1e50 overflows FP32: store 2^-s psi with
        // |2^-s psi|^2 ~ 1 at the mean density (an exact rescaling)
        c->psi_scale = (int)std::lround(0.5 * std::log2(std::max(cfg->rhoc, 1e-300) / c->ms));
        c->psc2 = std::ldexp(1.0, 2 * c->psi_scale);
        c->rho_mean = cfg->rhoc;
    }
    *out = c;
    return IFDM_OK;
}

void ifdm_destroy(ifdm_ctx *ctx) { release(ctx); }

int ifdm_set_psi(ifdm_ctx *c, const double *psi) {
    if (!c || !psi) return IFDM_E_ARG;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    c->pending = false;
    drop_checkpoints(c);
    if (!c->mixed) {
        HIPCHK(c, plane_copy(c, c->psi, const_cast<double *>(psi), 2 * sizeof(double), c->n, true,
                             hipMemcpyHostToDevice));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return IFDM_OK;
    }
    const size_t ne = 2 * (size_t)c->n3;
    std::unique_ptr<float[]> tmp(new (std::nothrow) float[ne]);
    if (!tmp) { c->err = "host staging buffer"; return IFDM_E_NOMEM; }
    double part[IFDM_HOST_CHUNKS];
    const int sc = c->psi_scale;
    host_chunks(ne, [&](size_t lo, size_t hi, int k) {
        double s = 0.0;
        for (size_t i = lo; i < hi; ++i) {
            tmp[i] = (float)std::ldexp(psi[i], -sc);
            s += psi[i] * psi[i];
        }
        part[k] = s;
    });
    double norm = 0.0;
    for (int k = 0; k < IFDM_HOST_CHUNKS; ++k) norm += part[k];
    c->rho_mean = c->ms * norm / (double)c->n3;
    HIPCHK(c, plane_copy(c, c->psi, tmp.get(), 2 * sizeof(float), c->n, true, hipMemcpyHostToDevice));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return IFDM_OK;
}

int ifdm_get_psi(ifdm_ctx *c, double *psi) {
    if (!c || !psi) return IFDM_E_ARG;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (int rc0 = ifdm_internal_flush(c)) return rc0;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (!c->mixed) {
        HIPCHK(c, plane_copy(c, c->psi, psi, 2 * sizeof(double), c->n, false, hipMemcpyDeviceToHost));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return IFDM_OK;
    }
    const size_t ne = 2 * (size_t)c->n3;
    std::unique_ptr<float[]> tmp(new (std::nothrow) float[ne]);
    if (!tmp) { c->err = "host staging buffer"; return IFDM_E_NOMEM; }
    HIPCHK(c, plane_copy(c, c->psi, tmp.get(), 2 * sizeof(float), c->n, false, hipMemcpyDeviceToHost));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int sc = c->psi_scale;
    host_chunks(ne, [&](size_t lo, size_t hi, int) {
        for (size_t i = lo; i < hi; ++i) psi[i] = std::ldexp((double)tmp[i], sc);
    });
    return IFDM_OK;
}

int ifdm_get_phi(ifdm_ctx *c, double *phi) {
    if (!c || !phi) return IFDM_E_ARG;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (!c->mixed) {
        HIPCHK(c, plane_copy(c, c->phi, phi, sizeof(double), c->n, false, hipMemcpyDeviceToHost));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return IFDM_OK;
    }
    const size_t ne = (size_t)c->n3;
    std::unique_ptr<float[]> tmp(new (std::nothrow) float[ne]);
    if (!tmp) { c->err = "host staging buffer"; return IFDM_E_NOMEM; }
    HIPCHK(c, plane_copy(c, c->phi, tmp.get(), sizeof(float), c->n, false, hipMemcpyDeviceToHost));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    host_chunks(ne, [&](size_t lo, size_t hi, int) {
        for (size_t i = lo; i < hi; ++i) phi[i] = tmp[i];
    });
    return IFDM_OK;
}

int ifdm_set_phi(ifdm_ctx *c, const double *phi) {
    if (!c || !phi) return IFDM_E_ARG;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (int rc0 = ifdm_internal_flush(c)) return rc0;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (!c->mixed) {
        HIPCHK(c, plane_copy(c, c->phi, const_cast<double *>(phi), sizeof(double), c->n, true, hipMemcpyHostToDevice));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return IFDM_OK;
    }
    const size_t ne = (size_t)c->n3;
    std::unique_ptr<float[]> tmp(new (std::nothrow) float[ne]);
    if (!tmp) { c->err = "host staging buffer"; return IFDM_E_NOMEM; }
    host_chunks(ne, [&](size_t lo, size_t hi, int) {
        for (size_t i = lo; i < hi; ++i) tmp[i] = (float)phi[i];
    });
    HIPCHK(c, plane_copy(c, c->phi, tmp.get(), sizeof(float), c->n, true, hipMemcpyHostToDevice));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return IFDM_OK;
}

int ifdm_poisson(ifdm_ctx *c, ifdm_stats *st) {
    if (!c) return IFDM_E_ARG;
    Dispatch d;
    get_dispatch(c->n, c->cfg.precision, &d);
    HIPCHK(c, hipSetDevice(c->cfg.device));
    int rc = d.poisson(c);
    if (rc) return rc;
    rc = read_stats(c, st);
    c->ev_pass.clear();  // a lone Poisson solve is not a profiled step
    return rc;
}

int ifdm_kick(ifdm_ctx *c, double dPE, double dQI) {
    if (!c) return IFDM_E_ARG;
    Dispatch d;
    get_dispatch(c->n, c->cfg.precision, &d);
    HIPCHK(c, hipSetDevice(c->cfg.device));
    drop_checkpoints(c);
    if (int rc0 = ifdm_internal_flush(c)) return rc0;
    int rc = d.kick(c, 0.5 * dPE * c->m_over_hbar, 0.5 * dQI * c->g_over_hbar);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return IFDM_OK;
}

int ifdm_drift(ifdm_ctx *c, double dKE) {
    if (!c) return IFDM_E_ARG;
    Dispatch d;
    get_dispatch(c->n, c->cfg.precision, &d);
    HIPCHK(c, hipSetDevice(c->cfg.device));
    drop_checkpoints(c);
    if (int rc0 = ifdm_internal_flush(c)) return rc0;
    int rc = d.drift(c, dKE);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return IFDM_OK;
}

int ifdm_fft3d(ifdm_ctx *c, int sign) {
    if (!c || (sign != -1 && sign != 1)) return IFDM_E_ARG;
    Dispatch d;
    get_dispatch(c->n, c->cfg.precision, &d);
    HIPCHK(c, hipSetDevice(c->cfg.device));
    drop_checkpoints(c);
    if (int rc0 = ifdm_internal_flush(c)) return rc0;
    int rc = d.fft3d(c, sign);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return IFDM_OK;
}

int ifdm_kdk_step(ifdm_ctx *c, const ifdm_dt *dt, ifdm_stats *st) {
    if (!c || !dt) return IFDM_E_ARG;
    Dispatch d;
    get_dispatch(c->n, c->cfg.precision, &d);
    HIPCHK(c, hipSetDevice(c->cfg.device));
    drop_checkpoints(c);
    if (int rc0 = ifdm_internal_flush(c)) return rc0;
    const double aphi = 0.5 * dt->dPE * c->m_over_hbar, arho = 0.5 * dt->dQI * c->g_over_hbar;
    int rc = d.kdp(c, aphi, arho, dt->dKE);
    if (rc) return rc;
    if ((rc = d.kick(c, aphi, arho))) return rc;
    if ((rc = read_stats(c, st))) return rc;
    return accumulate_profile(c);
}

int ifdm_step(ifdm_ctx *c, const ifdm_dt *dt, ifdm_stats *st) {
    if (!c || !dt) return IFDM_E_ARG;
    Dispatch d;
    get_dispatch(c->n, c->cfg.precision, &d);
    HIPCHK(c, hipSetDevice(c->cfg.device));
    drop_checkpoints(c);
    const double aphi = 0.5 * dt->dPE * c->m_over_hbar, arho = 0.5 * dt->dQI * c->g_over_hbar;
    double kphi = aphi, krho = arho;
    if (c->pending) {
        kphi += c->pend_aphi;
        krho += c->pend_arho;
    }
    int rc;
    c->run_stats_valid = false;
    if ((rc = d.kdp(c, kphi, krho, dt->dKE))) return rc;
    c->pending = true;
    c->pend_aphi = aphi;
    c->pend_arho = arho;
    if (st) {
        if ((rc = read_stats(c, st))) return rc;
        c->run_stats = *st;
        c->run_stats_valid = true;
    }
    return accumulate_profile(c);
}

int ifdm_internal_flush(ifdm_ctx *c) {
    if (!c->pending) return IFDM_OK;
    Dispatch d;
    get_dispatch(c->n, c->cfg.precision, &d);
    int rc;
    if ((rc = d.kick(c, c->pend_aphi, c->pend_arho))) return rc;
    c->pending = false;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return IFDM_OK;
}

int ifdm_flush(ifdm_ctx *c) {
    if (!c) return IFDM_E_ARG;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    return ifdm_internal_flush(c);
}

int ifdm_run(ifdm_ctx *c, const ifdm_dt *sched, int nsteps, ifdm_stats *st) {
    if (!c || (!sched && nsteps > 0) || nsteps < 0) return IFDM_E_ARG;
    int rc;
    for (int s = 0; s < nsteps; ++s)
        if ((rc = ifdm_step(c, &sched[s], st ? &st[s] : nullptr))) return rc;
    return ifdm_flush(c);
}

int ifdm_simulate_ex(ifdm_ctx *c, double a0, int nsteps, double *a_out, ifdm_dt *sched_out, int flags) {
    if (!c || nsteps < 0 || !(a0 > 0.0) || (flags & ~(IFDM_SIM_RESUME | IFDM_SIM_KEEP_PENDING))) return IFDM_E_ARG;
    int rc;
    ifdm_stats st;
    if (flags & IFDM_SIM_RESUME) {
        // inside a run: phi and the reductions are those of the last step,
        // whose second half-kick is still pending (it leaves |psi| alone)
        if (!c->pending || !c->run_stats_valid) {
            c->err = "ifdm_simulate_ex: IFDM_SIM_RESUME needs the pending half-kick of a previous step";
            return IFDM_E_STATE;
        }
        st = c->run_stats;
    } else if ((rc = ifdm_poisson(c, &st))) {
        return rc;                                               // :218 of the first step
    }
    double a = a0;
    if (c->dev_loop && nsteps > 0) {
        // the on-device loop (k_step_prep): the same steps as the host loop
        // below, bit for bit, with no host wait between steps
        Dispatch d;
        get_dispatch(c->n, c->cfg.precision, &d);
        HIPCHK(c, hipSetDevice(c->cfg.device));
        drop_checkpoints(c);
        StepIn in{};
        in.cc = crit_const(&c->cfg);
        in.a0 = a;
        in.max_phi = st.max_phi;
        in.max_rho = st.max_rho;
        in.norm = st.norm;
        in.has_pend = c->pending;
        in.pend_aphi = c->pend_aphi;
        in.pend_arho = c->pend_arho;
        in.psc2 = c->psc2;
        in.ms = c->ms;
        in.n3 = (double)c->n3;
        in.m_over_hbar = c->m_over_hbar;
        in.g_over_hbar = c->g_over_hbar;
        in.ke_coef = c->ke_coef;
        c->run_stats_valid = false;
        // steps between host waits: the record buffer, and with live
        // profiling the event pool (every pass of every step)
        int chunk = IFDM_REC_CAP;
        if (c->prof) chunk = std::min(chunk, (IFDM_EV_POOL / 2 - 1) / IFDM_NPASS);
        const double *rh = c->hrec + sizeof(DevStep) / sizeof(double);
        for (int s0 = 0; s0 < nsteps; s0 += chunk) {
            const int ns = std::min(chunk, nsteps - s0);
            for (int s = 0; s < ns; ++s) {
                in.first = s0 + s == 0;
                if ((rc = d.dev_step(c, in, c->drec + IFDM_REC * s))) return rc;
            }
            c->pending = true;
            HIPCHK(c, hipMemcpyAsync(c->hrec, c->dstep, sizeof(DevStep), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipMemcpyAsync((void *)rh, c->drec, sizeof(double) * IFDM_REC * ns, hipMemcpyDeviceToHost,
                                     c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            for (int s = 0; s < ns; ++s) {
                if (sched_out) {
                    sched_out[s0 + s].dKE = rh[IFDM_REC * s];
                    sched_out[s0 + s].dPE = rh[IFDM_REC * s + 1];
                    sched_out[s0 + s].dQI = rh[IFDM_REC * s + 2];
                }
                a = rh[IFDM_REC * s + 3];
            }
            if ((rc = accumulate_profile(c, ns))) return rc;
        }
        const DevStep ds = *(const DevStep *)c->hrec;
        c->pend_aphi = ds.pend_aphi;
        c->pend_arho = ds.pend_arho;
        // the last step's reductions, as ifdm_step reads them
        hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(256), 0, c->stream, c->red_stage, IFDM_RED_BLOCKS, c->red);
        LAUNCHCHK(c);
        if ((rc = read_stats(c, &st))) return rc;
        c->run_stats = st;
        c->run_stats_valid = true;
    }
    for (int s = 0; s < nsteps && !c->dev_loop; ++s) {
        ifdm_dt dt;
        double an;
        if ((rc = ifdm_timestep_criteria(&c->cfg, a, &st, &dt, &an, nullptr))) {
            c->err = "timestep criteria failed";
            return rc;
        }
        if (sched_out) sched_out[s] = dt;
        a = an;
        if ((rc = ifdm_step(c, &dt, &st))) return rc;
    }
    if (!(flags & IFDM_SIM_KEEP_PENDING) && (rc = ifdm_flush(c))) return rc;
    if (a_out) *a_out = a;
    return IFDM_OK;
}

int ifdm_simulate(ifdm_ctx *c, double a0, int nsteps, double *a_out, ifdm_dt *sched_out) {
    return ifdm_simulate_ex(c, a0, nsteps, a_out, sched_out, 0);
}

int ifdm_internal_theta_solve(ifdm_ctx *c, void *buf64, double coef) {
    Dispatch d;
    get_dispatch(c->n, IFDM_FP64, &d);
    if (!c->mixed && buf64 == c->psi) {
        int rc = d.theta_solve(c, coef);
        if (rc) return rc;
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return IFDM_OK;
    }
    // mixed mode: run the FP64 passes on a shadow context over buf64
    ifdm_ctx sh;
    sh.cfg = c->cfg;
    sh.cfg.precision = IFDM_FP64;
    sh.n = c->n;
    sh.n3 = c->n3;
    sh.ps = c->ps;
    sh.np = c->np;
    sh.hs = c->hs;
    sh.sps = c->sps;
    sh.stream = c->stream;
    sh.psi = buf64;
    sh.k2 = c->k2;
    sh.ke_coef = c->ke_coef;
    sh.green = c->green;
    sh.fftnorm = c->fftnorm;
    std::vector<double> tw;
    std::vector<float> twf;
    host_twiddles(c->n, tw, twf);
    HIPCHK(c, hipMalloc(&sh.tw, tw.size() * sizeof(double)));
    int rc = IFDM_OK;
    sh.tw64 = sh.tw;
    if (hipMemcpy(sh.tw, tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) rc = IFDM_E_HIP;
    if (!rc) rc = d.theta_solve(&sh, coef);
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = IFDM_E_HIP;
    if (rc && c->err.empty()) c->err = sh.err.empty() ? "theta solve failed" : sh.err;
    (void)hipFree(sh.tw);
    sh.stream = nullptr;
    return rc;
}

// ------------------------------------------------------------ diagnostics
int ifdm_power_spectrum(ifdm_ctx *c, int nbins, double *k_mean, double *pk, double *nmodes) {
    if (!c || nbins <= 0 || nbins > 4096 || !pk) return IFDM_E_ARG;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    Dispatch d;
    get_dispatch(c->n, c->cfg.precision, &d);
    double *dev = nullptr;
    HIPCHK(c, hipMalloc((void **)&dev, sizeof(double) * 3 * nbins));
    int rc = d.power_spectrum(c, nbins, dev, false);
    std::vector<double> h(3 * (size_t)nbins);
    ifdm_stats st;
    if (!rc) rc = read_stats(c, &st);
    if (!rc && hipMemcpy(h.data(), dev, sizeof(double) * h.size(), hipMemcpyDeviceToHost) != hipSuccess) {
        c->err = "power spectrum copy failed";
        rc = IFDM_E_HIP;
    }
    (void)hipFree(dev);
    c->ev_pass.clear();
    if (rc) return rc;
    // P(k) = L^3 |delta_k|^2 / N^6 with delta = (m_s|psi|^2 - rho_mean) / rho_mean
    const double L = 2.0 * c->cfg.xmax, n6 = std::pow((double)c->n, 6);
    const double rho_mean = c->ms * st.norm / (double)c->n3;
    const double kf = 2.0 * M_PI / L;
    for (int b = 0; b < nbins; ++b) {
        const double w = h[3 * b + 2];
        pk[b] = w > 0 ? std::pow(L, 3) * h[3 * b] / (rho_mean * rho_mean * n6) / w : 0.0;
        if (k_mean) k_mean[b] = w > 0 ? kf * h[3 * b + 1] / w : kf * (b + 1);
        if (nmodes) nmodes[b] = w;
    }
    return IFDM_OK;
}

int ifdm_power_spectrum_noalloc(ifdm_ctx *c, int nbins, double *k_mean, double *pk, double *nmodes) {
    if (!c || !pk) return IFDM_E_ARG;
    if (nbins <= 0 || nbins > IFDM_PK_MAX_BINS) {
        c->err = "ifdm_power_spectrum_noalloc: nbins must be in 1.." + std::to_string(IFDM_PK_MAX_BINS);
        return IFDM_E_ARG;
    }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    Dispatch d;
    get_dispatch(c->n, c->cfg.precision, &d);
    double *sums = c->pk_scratch, *outp = c->pk_scratch + 3 * IFDM_PK_MAX_BINS;
    int rc = d.power_spectrum(c, nbins, sums, true);
    c->ev_pass.clear();
    if (rc) return rc;
    const double L = 2.0 * c->cfg.xmax;
    hipLaunchKernelGGL(k_pk_finalize, dim3((nbins + 255) / 256), dim3(256), 0, c->stream, sums, nbins, c->red,
                       c->ms * c->psc2, (double)c->n3, std::pow(L, 3), std::pow((double)c->n, 6), 2.0 * M_PI / L,
                       outp);
    LAUNCHCHK(c);
    const size_t nb = sizeof(double) * (size_t)nbins;
    HIPCHK(c, hipMemcpyAsync(pk, outp, nb, hipMemcpyDeviceToHost, c->stream));
    if (k_mean) HIPCHK(c, hipMemcpyAsync(k_mean, outp + nbins, nb, hipMemcpyDeviceToHost, c->stream));
    if (nmodes) HIPCHK(c, hipMemcpyAsync(nmodes, outp + 2 * nbins, nb, hipMemcpyDeviceToHost, c->stream));
    // the reductions of this pass, as ifdm_power_spectrum leaves them
    return read_stats(c, nullptr);
}

// ---------------------------------------------------------------- adjoint
int ifdm_run_checkpointed(ifdm_ctx *c, const ifdm_dt *sched, int nsteps, int K) {
    if (!c || !sched || nsteps < 0 || K <= 0) return IFDM_E_ARG;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    int rc;
    if ((rc = ifdm_internal_flush(c))) return rc;
    const size_t bytes = (c->mixed ? 2 * sizeof(float) : 2 * sizeof(double)) * (size_t)c->np;
    if (c->ck) {
        (void)hipFree(c->ck);
        c->ck = nullptr;
    }
    c->ck_nsteps = -1;
    c->ck_sched.clear();
    const int nck = nsteps / K + 1;
    if (hipMalloc(&c->ck, bytes * nck) != hipSuccess) {
        c->ck = nullptr;
        c->err = "hipMalloc failed for the adjoint checkpoints";
        return IFDM_E_NOMEM;
    }
    // phi of the entry state: the run is a function of psi_0 alone, whatever
    // phi the context held before (the first half-kick reads it)
    {
        Dispatch d;
        get_dispatch(c->n, c->cfg.precision, &d);
        if ((rc = d.poisson(c))) return rc;
        c->ev_pass.clear();
    }
    HIPCHK(c, hipMemcpyAsync(c->ck, c->psi, bytes, hipMemcpyDeviceToDevice, c->stream));
    for (int s = 0; s < nsteps; ++s) {
        if ((rc = ifdm_step(c, &sched[s], nullptr))) return rc;
        if ((s + 1) % K == 0) {
            if ((rc = ifdm_internal_flush(c))) return rc;
            HIPCHK(c, hipMemcpyAsync((char *)c->ck + bytes * ((s + 1) / K), c->psi, bytes, hipMemcpyDeviceToDevice,
                                     c->stream));
        }
    }
    if ((rc = ifdm_internal_flush(c))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->ck_k = K;
    c->ck_count = nck;
    c->ck_nsteps = nsteps;
    c->ck_sched.assign(sched, sched + nsteps);
    return IFDM_OK;
}

// reverse sweep on the device gradient adj_g (see ifdm.h)
static int adjoint_sweep(ifdm_ctx *c, const ifdm_dt *sched, int nsteps) {
    Dispatch d;
    get_dispatch(c->n, c->cfg.precision, &d);
    const size_t bytes = (c->mixed ? 2 * sizeof(float) : 2 * sizeof(double)) * (size_t)c->np;
    int rc;
    c->ck_nsteps = -1;                                       // one sweep per forward run
    if ((rc = d.poisson(c))) return rc;                      // phi of the final state
    for (int n = nsteps - 1; n >= 0; --n) {
        const double aphi = 0.5 * sched[n].dPE * c->m_over_hbar, arho = 0.5 * sched[n].dQI * c->g_over_hbar;
        if ((rc = d.adj_kick(c, aphi, arho))) return rc;    // second half-kick
        if ((rc = d.drift(c, -sched[n].dKE))) return rc;    // psi back through the drift
        std::swap(c->psi, c->adj_g);
        rc = d.drift(c, -sched[n].dKE);                      // g <- D^H g
        std::swap(c->psi, c->adj_g);
        if (rc) return rc;
        if ((rc = d.poisson(c))) return rc;                  // phi at the start of step n
        if ((rc = d.adj_kick(c, aphi, arho))) return rc;    // first half-kick
        if (n % c->ck_k == 0) {                              // replace the reverse-evolved psi
            HIPCHK(c, hipMemcpyAsync(c->psi, (char *)c->ck + bytes * (n / c->ck_k), bytes, hipMemcpyDeviceToDevice,
                                     c->stream));
            if ((rc = d.poisson(c))) return rc;
        }
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return IFDM_OK;
}

// The checkpoints serve one reverse sweep: the sweep leaves psi_0 in psi
// (not psi_T), so a second sweep would start from the wrong state.
static int adjoint_prepare(ifdm_ctx *c, const ifdm_dt *sched, int nsteps) {
    if (!c->ck || c->ck_nsteps != nsteps) {
        c->err = "ifdm_adjoint needs the checkpoints of an ifdm_run_checkpointed over the same steps "
                 "(each run serves one reverse sweep)";
        return IFDM_E_STATE;
    }
    for (int s = 0; s < nsteps; ++s) {
        const ifdm_dt &a = c->ck_sched[s], &b = sched[s];
        if (a.dKE != b.dKE || a.dPE != b.dPE || a.dQI != b.dQI) {
            c->err = "ifdm_adjoint: schedule differs from the one of the last ifdm_run_checkpointed at step " +
                     std::to_string(s);
            return IFDM_E_STATE;
        }
    }
    const size_t es = c->mixed ? sizeof(float) : sizeof(double);
    const size_t cs = 2 * es * (size_t)c->np, rs = es * (size_t)c->np;
    if (!c->adj_g) {
        if (hipMalloc(&c->adj_g, cs) != hipSuccess || hipMalloc(&c->adj_s, rs) != hipSuccess ||
            hipMalloc(&c->adj_phs, rs) != hipSuccess) {
            c->err = "hipMalloc failed for the adjoint fields";
            return IFDM_E_NOMEM;
        }
        // the plane pads stay zero through the elementwise adjoint kernels
        HIPCHK(c, hipMemset(c->adj_g, 0, cs));
        HIPCHK(c, hipMemset(c->adj_s, 0, rs));
        HIPCHK(c, hipMemset(c->adj_phs, 0, rs));
    }
    return ifdm_internal_flush(c);
}

int ifdm_adjoint(ifdm_ctx *c, const ifdm_dt *sched, int nsteps, double *g) {
    if (!c || !sched || !g || nsteps < 0) return IFDM_E_ARG;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    int rc;
    if ((rc = adjoint_prepare(c, sched, nsteps))) return rc;
    const size_t n2 = 2 * (size_t)c->n3;
    if (!c->mixed) {
        HIPCHK(c, plane_copy(c, c->adj_g, g, 2 * sizeof(double), c->n, true, hipMemcpyHostToDevice));
        if ((rc = adjoint_sweep(c, sched, nsteps))) return rc;
        HIPCHK(c, plane_copy(c, c->adj_g, g, 2 * sizeof(double), c->n, false, hipMemcpyDeviceToHost));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return IFDM_OK;
    }
    // mixed: the sweep is linear in g, so g is held as 2^-sg g in FP32 with
    // |2^-sg g| <= 1 at the start (the psi scale is folded into the kicks)
    double gpart[IFDM_HOST_CHUNKS];
    host_chunks(n2, [&](size_t lo, size_t hi, int k) {
        double m = 0.0;
        for (size_t i = lo; i < hi; ++i) m = std::max(m, std::fabs(g[i]));
        gpart[k] = m;
    });
    double gmax = 0.0;
    for (int k = 0; k < IFDM_HOST_CHUNKS; ++k) gmax = std::max(gmax, gpart[k]);
    const int sg = gmax > 0 ? std::ilogb(gmax) + 1 : 0;
    std::unique_ptr<float[]> tmp(new (std::nothrow) float[n2]);
    if (!tmp) { c->err = "host staging buffer"; return IFDM_E_NOMEM; }
    host_chunks(n2, [&](size_t lo, size_t hi, int) {
        for (size_t i = lo; i < hi; ++i) tmp[i] = (float)std::ldexp(g[i], -sg);
    });
    HIPCHK(c, plane_copy(c, c->adj_g, tmp.get(), 2 * sizeof(float), c->n, true, hipMemcpyHostToDevice));
    if ((rc = adjoint_sweep(c, sched, nsteps))) return rc;
    HIPCHK(c, plane_copy(c, c->adj_g, tmp.get(), 2 * sizeof(float), c->n, false, hipMemcpyDeviceToHost));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // back to physical units: g scale 2^sg, and the psi scale of the kicked
    // field enters g = e^{i theta} g' + 2 psi A s linearly through psi and s
    host_chunks(n2, [&](size_t lo, size_t hi, int) {
        for (size_t i = lo; i < hi; ++i) g[i] = std::ldexp((double)tmp[i], sg);
    });
    return IFDM_OK;
}

int ifdm_adjoint_device(ifdm_ctx *c, const ifdm_dt *sched, int nsteps, void *g_dev) {
    if (!c || !sched || !g_dev || nsteps < 0) return IFDM_E_ARG;
    if (c->mixed) { c->err = "ifdm_adjoint_device runs in FP64 contexts (use ifdm_adjoint)"; return IFDM_E_STATE; }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    int rc;
    if ((rc = adjoint_prepare(c, sched, nsteps))) return rc;
    HIPCHK(c, plane_copy(c, c->adj_g, g_dev, 2 * sizeof(double), c->n, true, hipMemcpyDeviceToDevice));
    if ((rc = adjoint_sweep(c, sched, nsteps))) return rc;
    HIPCHK(c, plane_copy(c, c->adj_g, g_dev, 2 * sizeof(double), c->n, false, hipMemcpyDeviceToDevice));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return IFDM_OK;
}

int ifdm_set_psi_device(ifdm_ctx *c, const void *psi_dev) {
    if (!c || !psi_dev) return IFDM_E_ARG;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    c->pending = false;
    drop_checkpoints(c);
    const long n2 = (long)c->n * c->n;
    if (!c->mixed) {
        HIPCHK(c, plane_copy(c, c->psi, const_cast<void *>(psi_dev), 2 * sizeof(double), c->n, true,
                             hipMemcpyDeviceToDevice));
    } else {
        const unsigned nb = (unsigned)std::min<long>((c->n3 + 255) / 256, 65536);
        hipLaunchKernelGGL((k_store_scaled<float>), dim3(nb), dim3(256), 0, c->stream, (const cx<double> *)psi_dev,
                           (cx<float> *)c->psi, c->n3, n2, n2, c->ps, c->psi_scale);
        LAUNCHCHK(c);
        // the FP64 mean density m_s <|psi|^2> that the complex-float Poisson
        // source subtracts, from the FP64 input (as ifdm_set_psi does)
        hipLaunchKernelGGL(k_norm_partial, dim3(IFDM_RED_BLOCKS), dim3(256), 0, c->stream,
                           (const cx<double> *)psi_dev, c->n3, c->red_stage);
        LAUNCHCHK(c);
        std::vector<double> part(IFDM_RED_BLOCKS);
        HIPCHK(c, hipMemcpyAsync(part.data(), c->red_stage, sizeof(double) * IFDM_RED_BLOCKS, hipMemcpyDeviceToHost,
                                 c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        double norm = 0.0;
        for (double x : part) norm += x;
        c->rho_mean = c->ms * norm / (double)c->n3;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return IFDM_OK;
}

int ifdm_get_psi_device(ifdm_ctx *c, void *psi_dev) {
    if (!c || !psi_dev) return IFDM_E_ARG;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (int rc0 = ifdm_internal_flush(c)) return rc0;
    const long n2 = (long)c->n * c->n;
    if (!c->mixed) {
        HIPCHK(c, plane_copy(c, c->psi, psi_dev, 2 * sizeof(double), c->n, false, hipMemcpyDeviceToDevice));
    } else {
        const unsigned nb = (unsigned)std::min<long>((c->n3 + 255) / 256, 65536);
        hipLaunchKernelGGL(k_load_scaled, dim3(nb), dim3(256), 0, c->stream, (const cx<float> *)c->psi,
                           (cx<double> *)psi_dev, c->n3, n2, c->ps, n2, c->psi_scale);
        LAUNCHCHK(c);
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return IFDM_OK;
}

int ifdm_stream_copy(ifdm_ctx *c, size_t bytes, int iters, double *gbs) {
    if (!c || !gbs || iters <= 0 || bytes < 16) return IFDM_E_ARG;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const long n = (long)(bytes / 16);
    void *a = nullptr, *b = nullptr;
    HIPCHK(c, hipMalloc(&a, n * 16));
    if (hipMalloc(&b, n * 16) != hipSuccess) {
        (void)hipFree(a);
        c->err = "hipMalloc failed in ifdm_stream_copy";
        return IFDM_E_NOMEM;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = IFDM_OK;
    float ms = 0.f;
    const unsigned nb = (unsigned)std::min<long>((n + 255) / 256, 1L << 20);
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
        hipMemsetAsync(a, 0, n * 16, c->stream) != hipSuccess) {
        c->err = "event or memset failed in ifdm_stream_copy";
        rc = IFDM_E_HIP;
    } else {
        hipLaunchKernelGGL(k_copy, dim3(nb), dim3(256), 0, c->stream, (const double2 *)a, (double2 *)b, n);  // warm
        if (hipEventRecord(e0, c->stream) != hipSuccess) rc = IFDM_E_HIP;
        for (int i = 0; i < iters && !rc; ++i)
            hipLaunchKernelGGL(k_copy, dim3(nb), dim3(256), 0, c->stream, (const double2 *)a, (double2 *)b, n);
        if (!rc && (hipGetLastError() != hipSuccess || hipEventRecord(e1, c->stream) != hipSuccess ||
                    hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = IFDM_E_HIP;
        if (rc) c->err = "copy kernel failed";
    }
    if (!rc) *gbs = 2.0 * 16.0 * (double)n * iters / (ms * 1e-3) / 1e9;
    // release whatever was created (errors here leave nothing to undo)
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}

int ifdm_set_option(ifdm_ctx *c, const char *name, long value) {
    if (!c || !name) return IFDM_E_ARG;
    if (std::strcmp(name, "persistent") == 0) {
        if (value < 0 || value > 31) { c->err = "persistent must be a mask in 0..31"; return IFDM_E_ARG; }
        c->persistent = (int)value;
        return IFDM_OK;
    }
    if (std::strcmp(name, "profile_mask") == 0) {
        if (value <= 0 || value >= (1L << IFDM_NPASS)) {
            c->err = "profile_mask must be a non-empty mask of the " + std::to_string(IFDM_NPASS) + " passes";
            return IFDM_E_ARG;
        }
        c->prof_mask = (unsigned)value;
        return IFDM_OK;
    }
    if (std::strcmp(name, "device_loop") == 0) {
        if (value != 0 && value != 1) { c->err = "device_loop must be 0 (time step on the host) or 1 (on the device)"; return IFDM_E_ARG; }
        c->dev_loop = (int)value;
        return IFDM_OK;
    }
    if (std::strcmp(name, "ic_deposit") == 0) {
        if (value != 0 && value != 1) { c->err = "ic_deposit must be 0 (reference order) or 1 (FP64 atomics)"; return IFDM_E_ARG; }
        if (c->ic_rho) { c->err = "ic_deposit cannot change between ifdm_ic_begin and ifdm_ic_finish"; return IFDM_E_STATE; }
        c->ic_mode = (int)value;
        return IFDM_OK;
    }
    c->err = std::string("unknown option ") + name;
    return IFDM_E_ARG;
}

int ifdm_set_profiling(ifdm_ctx *c, int on) {
    if (!c) return IFDM_E_ARG;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (on && c->ev_pool.empty()) {
        c->ev_pool.resize(IFDM_EV_POOL, nullptr);
        // timing-only events: no system-scope fence (cache write-back and
        // invalidation) when they complete, which would slow the launch after
        // each of them and lengthen the times they bracket
        for (auto &e : c->ev_pool) HIPCHK(c, hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    }
    c->prof = on != 0;
    c->ev_pass.clear();
    for (double &x : c->pass_ms) x = 0.0;
    c->pass_count = 0;
    return IFDM_OK;
}

int ifdm_get_profile(const ifdm_ctx *c, double *ms_sum, long *count) {
    if (!c || !ms_sum || !count) return IFDM_E_ARG;
    for (int p = 0; p < IFDM_NPASS; ++p) ms_sum[p] = c->pass_ms[p];
    *count = c->pass_count;
    return IFDM_OK;
}

int ifdm_time_passes(ifdm_ctx *c, const ifdm_dt *dt, int iters, float *ms_out) {
    if (!c || !dt || iters <= 0 || !ms_out) return IFDM_E_ARG;
    Dispatch d;
    get_dispatch(c->n, c->cfg.precision, &d);
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const double aphi = 0.5 * dt->dPE * c->m_over_hbar, arho = 0.5 * dt->dQI * c->g_over_hbar;
    hipEvent_t ev[IFDM_NPASS + 1];
    for (auto &e : ev) HIPCHK(c, hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    std::vector<double> acc(IFDM_NPASS, 0.0);
    int rc = 0;
    for (int it = 0; it < iters && !rc; ++it) {
        for (int p = 0; p < IFDM_NPASS && !rc; ++p) {
            if (hipEventRecord(ev[p], c->stream) != hipSuccess) { rc = IFDM_E_HIP; break; }
            rc = d.pass(c, p, 2 * aphi, 2 * arho, dt->dKE);
            if (!rc && hipEventRecord(ev[p + 1], c->stream) != hipSuccess) rc = IFDM_E_HIP;
        }
        if (!rc && hipEventSynchronize(ev[IFDM_NPASS]) != hipSuccess) rc = IFDM_E_HIP;
        for (int p = 0; p < IFDM_NPASS && !rc; ++p) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, ev[p], ev[p + 1]) != hipSuccess) { rc = IFDM_E_HIP; break; }
            acc[p] += ms;
        }
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    if (rc) { if (c->err.empty()) c->err = "event timing failed"; return rc; }
    for (int p = 0; p < IFDM_NPASS; ++p) ms_out[p] = (float)(acc[p] / iters);
    return IFDM_OK;
}

}  // extern "C"
