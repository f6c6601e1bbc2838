// ifdm_io.hip -- initial conditions and snapshots of libifdm.so:
// the Gadget-1 reader (src/gadget_IC.h:259-353, 573-672), the device
// cloud-in-cell deposit and phase assembly (ic_kernels.h), and the
// output<t>.dat writer / rerun reader (src/write3D.h:7-97,
// src/gadget_IC.h:709-750) with 64-bit offsets throughout.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ifdm.h"
#include "constants.h"
#include "ctx.h"
#include "ic_kernels.h"

#include <rocprim/device/device_radix_sort.hpp>

using namespace ifdm;

namespace {

unsigned grid_for(long n) { return (unsigned)std::min<long>((n + 255) / 256, 1L << 16); }

void free_ic(ifdm_ctx *c) {
    void *b[] = {c->ic_rho, c->ic_vx, c->ic_vy, c->ic_vz, c->ic_skipped, c->ic_first, c->ic_last};
    for (void *p : b)
        if (p) (void)hipFree(p);
    c->ic_rho = nullptr;
    c->ic_vx = c->ic_vy = c->ic_vz = nullptr;
    c->ic_first = c->ic_last = nullptr;
    c->ic_skipped = nullptr;
}

// Device buffers that live for one call; freed on every return path.
struct DevBufs {
    std::vector<void *> p;
    ~DevBufs() {
        for (void *q : p)
            if (q) (void)hipFree(q);
    }
    template <typename X> bool get(X **out, size_t bytes) {
        void *q = nullptr;
        if (hipMalloc(&q, bytes ? bytes : 1) != hipSuccess) return false;
        p.push_back(q);
        *out = (X *)q;
        return true;
    }
};

// The reference-order deposit of one ifdm_ic_deposit call: the sequence of
// ng ghost copies then np particles (pos, vel on the device), in batches of
// at most 2^27 entries, each sorted by base cell and gathered per target cell.
int deposit_ordered(ifdm_ctx *c, const float *dpos, const float *dvel, const CicOrd &a) {
    const long long total = a.ng + a.np;
    if (total == 0) return IFDM_OK;
    if (total >= (1ll << 32) - 1) { c->err = "too many particles in one ifdm_ic_deposit call"; return IFDM_E_ARG; }
    const unsigned n3 = (unsigned)c->n3;
    const long long batch = std::min<long long>(total, 1ll << 27);
    unsigned end_bit = 1;
    while ((1ull << end_bit) <= (unsigned long long)n3) ++end_bit;   // keys <= n3
    DevBufs B;
    unsigned *kin, *kout, *vin, *vout;
    if (!B.get(&kin, 4 * batch) || !B.get(&kout, 4 * batch) || !B.get(&vin, 4 * batch) || !B.get(&vout, 4 * batch)) {
        c->err = "hipMalloc failed for the deposit's sort buffers";
        return IFDM_E_NOMEM;
    }
    size_t tmp_bytes = 0;
    HIPCHK(c, rocprim::radix_sort_pairs(nullptr, tmp_bytes, kin, kout, vin, vout, (size_t)batch, 0, end_bit, c->stream));
    void *tmp;
    if (!B.get(&tmp, tmp_bytes)) { c->err = "hipMalloc failed for the deposit's sort storage"; return IFDM_E_NOMEM; }
    for (long long s0 = 0; s0 < total; s0 += batch) {
        const long long m = std::min(batch, total - s0);
        const unsigned nb = (unsigned)((m + 255) / 256);
        hipLaunchKernelGGL(k_cico_keys, dim3(nb), dim3(256), 0, c->stream, dpos, a, s0, m, kin, vin, c->ic_skipped);
        LAUNCHCHK(c);
        size_t tb = tmp_bytes;
        HIPCHK(c, rocprim::radix_sort_pairs(tmp, tb, kin, kout, vin, vout, (size_t)m, 0, end_bit, c->stream));
        HIPCHK(c, hipMemsetAsync(c->ic_first, 0, sizeof(unsigned) * n3, c->stream));
        HIPCHK(c, hipMemsetAsync(c->ic_last, 0, sizeof(unsigned) * n3, c->stream));
        hipLaunchKernelGGL(k_cico_bounds, dim3(nb), dim3(256), 0, c->stream, kout, m, n3, c->ic_first, c->ic_last);
        LAUNCHCHK(c);
        hipLaunchKernelGGL((k_cico_gather<float>), dim3(grid_for(c->n3)), dim3(256), 0, c->stream, dpos, dvel, a, vout,
                           c->ic_first, c->ic_last, c->ic_rho, (float *)c->ic_vx, (float *)c->ic_vy,
                           (float *)c->ic_vz);
        LAUNCHCHK(c);
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return IFDM_OK;
}

struct File {
    FILE *f = nullptr;
    File(const std::string &p, const char *mode) { f = std::fopen(p.c_str(), mode); }
    ~File() { if (f) std::fclose(f); }
};

// Gadget-1 header fields used by the reference (src/gadget_IC.h:17-33)
struct GadgetHeader {
    int npart[6];
    double mass[6];
    double time;
    double box;
    double hubble;
};

bool parse_header(const unsigned char *b, GadgetHeader *h) {
    std::memcpy(h->npart, b + 0, 24);
    std::memcpy(h->mass, b + 24, 48);
    std::memcpy(&h->time, b + 72, 8);
    std::memcpy(&h->box, b + 128, 8);
    std::memcpy(&h->hubble, b + 152, 8);
    return h->npart[1] >= 0;
}

}  // namespace

extern "C" {

int ifdm_ic_begin(ifdm_ctx *c, double box, double pmass) {
    if (!c || !(box > 0) || !(pmass > 0)) return IFDM_E_ARG;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    free_ic(c);
    c->ic_f32 = c->ic_mode == 0;
    const size_t n3 = (size_t)c->n3, vb = (c->ic_f32 ? sizeof(float) : sizeof(double)) * n3;
    struct { void **p; size_t bytes; } g[] = {{(void **)&c->ic_rho, sizeof(double) * n3},
                                              {&c->ic_vx, vb}, {&c->ic_vy, vb}, {&c->ic_vz, vb},
                                              {(void **)&c->ic_first, c->ic_f32 ? sizeof(unsigned) * n3 : 0},
                                              {(void **)&c->ic_last, c->ic_f32 ? sizeof(unsigned) * n3 : 0}};
    for (auto &b : g) {
        if (!b.bytes) continue;
        if (hipMalloc(b.p, b.bytes) != hipSuccess) {
            free_ic(c);
            c->err = "hipMalloc failed for the initial-condition grids";
            return IFDM_E_NOMEM;
        }
        HIPCHK(c, hipMemsetAsync(*b.p, 0, b.bytes, c->stream));
    }
    HIPCHK(c, hipMalloc((void **)&c->ic_skipped, sizeof(unsigned long long)));
    HIPCHK(c, hipMemsetAsync(c->ic_skipped, 0, sizeof(unsigned long long), c->stream));
    c->ic_box = box;
    c->ic_pmass = pmass;
    c->ic_npart = 0;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return IFDM_OK;
}

int ifdm_ic_deposit(ifdm_ctx *c, const float *pos, const float *vel, long long np, int ref_ghosts) {
    if (!c || (np > 0 && (!pos || !vel)) || np < 0) return IFDM_E_ARG;
    if (!c->ic_rho) { c->err = "ifdm_ic_deposit before ifdm_ic_begin"; return IFDM_E_STATE; }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    CicArgs a;
    a.n = c->n;
    a.box = c->ic_box;
    a.lg = c->ic_box / c->n;
    a.pmass = c->ic_pmass;
    a.l3 = std::pow(a.lg, 3);
    // the reference's one-rank "left ghosts": trailing particles of the
    // x-sorted chunk beyond the last cell centre are deposited once more
    // (src/gadget_IC.h:445-448)
    long long ghost0 = np;
    if (ref_ghosts) {
        const double xlast = (c->n - 1) * a.lg + 0.5 * a.lg;
        for (long long pn = np - 1; pn > 0; --pn) {
            if (xlast < pos[3 * pn]) ghost0 = pn;
            else break;
        }
    }
    if (c->ic_f32) {
        // reference order: the whole call's particles on the device (the
        // ghosts index into the chunk's end)
        DevBufs B;
        float *dp, *dv;
        if (!B.get(&dp, sizeof(float) * 3 * (size_t)np) || !B.get(&dv, sizeof(float) * 3 * (size_t)np)) {
            c->err = "hipMalloc failed for the particles";
            return IFDM_E_NOMEM;
        }
        HIPCHK(c, hipMemcpy(dp, pos, sizeof(float) * 3 * (size_t)np, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(dv, vel, sizeof(float) * 3 * (size_t)np, hipMemcpyHostToDevice));
        std::vector<double> xg(c->n);
        for (int i = 0; i < c->n; ++i) xg[i] = i * a.lg + 0.5 * a.lg;                  // src/gadget_IC.h:333
        double *dxg;
        if (!B.get(&dxg, sizeof(double) * c->n)) { c->err = "hipMalloc failed"; return IFDM_E_NOMEM; }
        HIPCHK(c, hipMemcpy(dxg, xg.data(), sizeof(double) * c->n, hipMemcpyHostToDevice));
        CicOrd o;
        o.n = c->n;
        o.box = a.box;
        o.lg = a.lg;
        o.hl = 0.5 * a.lg;
        o.pmass = a.pmass;
        o.l3 = a.l3;
        o.xg = dxg;
        o.np = np;
        o.ng = np - ghost0;
        const int rc = deposit_ordered(c, dp, dv, o);
        if (!rc) c->ic_npart += np + (np - ghost0);
        return rc;
    }
    const long long batch = 1ll << 24;
    float *dp = nullptr, *dv = nullptr;
    const long long cap = std::min<long long>(std::max<long long>(np, 1), batch);
    HIPCHK(c, hipMalloc((void **)&dp, sizeof(float) * 3 * cap));
    if (hipMalloc((void **)&dv, sizeof(float) * 3 * cap) != hipSuccess) {
        (void)hipFree(dp);
        c->err = "hipMalloc failed for a particle batch";
        return IFDM_E_NOMEM;
    }
    int rc = IFDM_OK;
    auto deposit = [&](long long p0, long long p1) {
        for (long long b = p0; b < p1 && !rc; b += batch) {
            const long long m = std::min(batch, p1 - b);
            if (hipMemcpy(dp, pos + 3 * b, sizeof(float) * 3 * m, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(dv, vel + 3 * b, sizeof(float) * 3 * m, hipMemcpyHostToDevice) != hipSuccess) {
                rc = IFDM_E_HIP;
                c->err = "particle upload failed";
                break;
            }
            hipLaunchKernelGGL(k_cic, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, c->stream, dp, dv, (long)m, a,
                               c->ic_rho, (double *)c->ic_vx, (double *)c->ic_vy, (double *)c->ic_vz,
                               c->ic_skipped);
            if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) {
                rc = IFDM_E_HIP;
                c->err = "CIC deposit kernel failed";
            }
        }
    };
    if (ghost0 < np) deposit(ghost0, np);
    deposit(0, np);
    (void)hipFree(dp);
    (void)hipFree(dv);
    if (!rc) c->ic_npart += np + (np - ghost0);
    return rc;
}

int ifdm_ic_finish(ifdm_ctx *c, double a, double *theta_out, float *vx_out, long long *skipped_out) {
    if (!c || !(a > 0)) return IFDM_E_ARG;
    if (!c->ic_rho) { c->err = "ifdm_ic_finish before ifdm_ic_begin"; return IFDM_E_STATE; }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const long n3 = c->n3;
    c->pending = false;
    c->ck_nsteps = -1;   // psi replaced: the adjoint checkpoints no longer apply
    if (c->ic_f32)
        hipLaunchKernelGGL((k_ic_velnorm<float>), dim3(grid_for(n3)), dim3(256), 0, c->stream, c->ic_rho,
                           (float *)c->ic_vx, (float *)c->ic_vy, (float *)c->ic_vz, n3, c->ic_pmass);
    else
        hipLaunchKernelGGL((k_ic_velnorm<double>), dim3(grid_for(n3)), dim3(256), 0, c->stream, c->ic_rho,
                           (double *)c->ic_vx, (double *)c->ic_vy, (double *)c->ic_vz, n3, c->ic_pmass);
    HIPCHK(c, hipGetLastError());
    if (vx_out) {
        // write_binary_vg writes the float grid after the normalisation
        // (src/gadget_IC.h:656, src/write3D.h:24-53)
        if (c->ic_f32) {
            HIPCHK(c, hipMemcpyAsync(vx_out, c->ic_vx, sizeof(float) * n3, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
        } else {
            std::vector<double> tmp((size_t)n3);
            HIPCHK(c, hipMemcpyAsync(tmp.data(), c->ic_vx, sizeof(double) * n3, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            for (long i = 0; i < n3; ++i) vx_out[i] = (float)tmp[i];
        }
    }
    // theta work buffer (psi's plane stride): psi itself in FP64 mode
    void *buf = c->psi;
    if (c->mixed) {
        HIPCHK(c, hipMalloc(&buf, 2 * sizeof(double) * (size_t)c->np));
        HIPCHK(c, hipMemsetAsync(buf, 0, 2 * sizeof(double) * (size_t)c->np, c->stream));
    }
    const double dx = 2.0 * c->cfg.xmax / c->n;                          // par.dx, src/struct3D.h:52
    const double unit = 1e3 / m2Mpc * s2Myr;                              // src/gadget_IC.h:171
    if (c->ic_f32)
        hipLaunchKernelGGL((k_ic_divergence<float>), dim3(grid_for(n3)), dim3(256), 0, c->stream,
                           (const float *)c->ic_vx, (const float *)c->ic_vy, (const float *)c->ic_vz,
                           (cx<double> *)buf, c->n, c->ps, unit, dx);
    else
        hipLaunchKernelGGL((k_ic_divergence<double>), dim3(grid_for(n3)), dim3(256), 0, c->stream,
                           (const double *)c->ic_vx, (const double *)c->ic_vy, (const double *)c->ic_vz,
                           (cx<double> *)buf, c->n, c->ps, unit, dx);
    int rc = hipGetLastError() == hipSuccess ? IFDM_OK : IFDM_E_HIP;
    if (!rc) rc = ifdm_internal_theta_solve(c, buf, -std::pow(a, 3. / 2));   // src/gadget_IC.h:233
    if (!rc && theta_out) {
        if (plane_copy(c, buf, theta_out, 2 * sizeof(double), c->n, false, hipMemcpyDeviceToHost) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            rc = IFDM_E_HIP;
    }
    if (!rc) {
        const double uc = std::pow(h_0, 2) * std::pow(1e3, 3);            // src/gadget_IC.h:659
        const double cph = c->cfg.mfdm / std::pow(c_light, 2) / hbar;
        hipLaunchKernelGGL(k_ic_assemble, dim3(grid_for(n3)), dim3(256), 0, c->stream, c->ic_rho, (cx<double> *)buf,
                           c->n, c->ps, uc, c->ms, cph);
        if (hipGetLastError() != hipSuccess) rc = IFDM_E_HIP;
    }
    if (!rc && c->mixed) {
        hipLaunchKernelGGL((k_store_scaled<float>), dim3(grid_for(n3)), dim3(256), 0, c->stream,
                           (const cx<double> *)buf, (cx<float> *)c->psi, n3, (long)c->n * c->n, c->ps, c->ps,
                           c->psi_scale);
        if (hipGetLastError() != hipSuccess) rc = IFDM_E_HIP;
    }
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = IFDM_E_HIP;
    if (skipped_out && !rc) {
        unsigned long long sk = 0;
        if (hipMemcpy(&sk, c->ic_skipped, sizeof(sk), hipMemcpyDeviceToHost) != hipSuccess) rc = IFDM_E_HIP;
        *skipped_out = (long long)sk;
    }
    if (c->mixed && buf) (void)hipFree(buf);
    free_ic(c);
    if (rc && c->err.empty()) c->err = "initial-condition assembly failed";
    return rc;
}

int ifdm_read_gadget(ifdm_ctx *c, const char *icpath, int nchunks, int ref_ghosts, const char *voutput_path,
                     ifdm_ic_info *info) {
    if (!c || !icpath || nchunks <= 0) return IFDM_E_ARG;
    double a = 0.0, rhoc = 0.0, box = 0.0, pmass = 0.0;
    long long total = 0;
    int rc;
    for (int chunk = 0; chunk < nchunks; ++chunk) {
        const std::string path = std::string(icpath) + "." + std::to_string(chunk);      // :279
        File f(path, "rb");
        if (!f.f) { c->err = "cannot read IC file " + path; return IFDM_E_IO; }
        int32_t m0 = 0, m1 = 0;
        unsigned char hb[256];
        if (std::fread(&m0, 4, 1, f.f) != 1 || std::fread(hb, 1, 256, f.f) != 256 || std::fread(&m1, 4, 1, f.f) != 1) {
            c->err = "truncated Gadget header in " + path;
            return IFDM_E_FORMAT;
        }
        GadgetHeader h;
        parse_header(hb, &h);
        if (m0 != 256 || m1 != 256 || h.npart[1] < 0) { c->err = "not a Gadget-1 file: " + path; return IFDM_E_FORMAT; }
        // box check (src/gadget_IC.h:309-314): BoxSize [kpc/h] == 2 xmax 1e3 h_0
        if (!(std::fabs(h.box - 2 * c->cfg.xmax * 1e3 * h_0) < 0.001)) {
            c->err = "Boxsize of IC is different to xmax in Simulation: " + std::to_string(2 * c->cfg.xmax * 1e3 * h_0) +
                     " vs " + std::to_string(h.box);
            return IFDM_E_FORMAT;
        }
        const long long np = h.npart[1];
        const double mass = h.mass[1] * 1e10;                                             // :317
        a = h.time;                                                                       // :318
        rhoc = nchunks * (double)h.npart[1] * mass / std::pow(h.box, 3) * std::pow(1e3, 3) * std::pow(h_0, 2);  // :319
        if (chunk == 0) {
            box = h.box;
            pmass = mass;
            if ((rc = ifdm_ic_begin(c, box, pmass))) return rc;
        } else if (h.box != box || mass != pmass) {
            c->err = "IC chunks disagree on BoxSize or particle mass";
            return IFDM_E_FORMAT;
        }
        std::vector<float> pos(3 * (size_t)np), vel(3 * (size_t)np);
        int32_t b0 = 0, b1 = 0, b2 = 0, b3 = 0;
        const size_t want = 3 * (size_t)np;
        if (std::fread(&b0, 4, 1, f.f) != 1 || std::fread(pos.data(), sizeof(float), want, f.f) != want ||
            std::fread(&b1, 4, 1, f.f) != 1 || std::fread(&b2, 4, 1, f.f) != 1 ||
            std::fread(vel.data(), sizeof(float), want, f.f) != want || std::fread(&b3, 4, 1, f.f) != 1) {
            c->err = "truncated particle blocks in " + path;
            return IFDM_E_FORMAT;
        }
        if ((rc = ifdm_ic_deposit(c, pos.data(), vel.data(), np, ref_ghosts))) return rc;
        total += np;
    }
    long long skipped = 0;
    std::vector<float> vx;
    if (voutput_path) vx.resize((size_t)c->n3);
    if ((rc = ifdm_ic_finish(c, a, nullptr, voutput_path ? vx.data() : nullptr, &skipped))) return rc;
    c->cfg.rhoc = rhoc;                                                                   // par.rhoc
    if (c->mixed) c->rho_mean = rhoc;
    if (voutput_path) {
        // write_binary_vg (src/write3D.h:24-53): the vx grid as float, appended
        File f(voutput_path, "ab");
        if (!f.f || std::fwrite(vx.data(), sizeof(float), vx.size(), f.f) != vx.size()) {
            c->err = std::string("cannot write ") + voutput_path;
            return IFDM_E_IO;
        }
    }
    if (info) {
        info->box = box;
        info->pmass = pmass;
        info->a = a;
        info->rhoc = rhoc;
        info->npart = total;
        info->skipped = skipped;
    }
    return IFDM_OK;
}

int ifdm_set_rhoc(ifdm_ctx *c, double rhoc) {
    if (!c || !(rhoc >= 0)) return IFDM_E_ARG;
    c->cfg.rhoc = rhoc;
    return IFDM_OK;
}

// output<t>.dat: [f64 mfdm eV][f64 g eV cm^3][f64 a][f64 x[N]][c128 psi N^3]
// (write_binary_params + write_binary_wfc, src/write3D.h:7-23, 55-77)
int ifdm_write_snapshot(ifdm_ctx *c, const char *path, double a, double mfdm_ev, double g_ev_cm3) {
    if (!c || !path) return IFDM_E_ARG;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    int rc;
    if ((rc = ifdm_internal_flush(c))) return rc;
    File f(path, "wb");
    if (!f.f) { c->err = std::string("cannot open ") + path; return IFDM_E_IO; }
    const int n = c->n;
    const double xmax = c->cfg.xmax;
    std::vector<double> x(n);
    for (int i = 0; i < n; ++i) x[i] = xmax / n - xmax + i * (2.0 * xmax / n);         // src/struct3D.h:72
    if (std::fwrite(&mfdm_ev, 8, 1, f.f) != 1 || std::fwrite(&g_ev_cm3, 8, 1, f.f) != 1 ||
        std::fwrite(&a, 8, 1, f.f) != 1 || std::fwrite(x.data(), 8, n, f.f) != (size_t)n) {
        c->err = std::string("write failed: ") + path;
        return IFDM_E_IO;
    }
    // stream psi plane by plane (64-bit sizes: 1024^3 complex128 is 17 GB)
    const size_t plane = (size_t)n * n;
    const size_t planes_per = std::max<size_t>(1, (size_t(1) << 27) / plane);
    std::vector<double> buf(2 * plane * planes_per);
    std::vector<float> fbuf(c->mixed ? 2 * plane * planes_per : 0);
    for (size_t p0 = 0; p0 < (size_t)n; p0 += planes_per) {
        const size_t np = std::min(planes_per, (size_t)n - p0), cnt = 2 * plane * np;
        if (!c->mixed) {
            HIPCHK(c, plane_copy(c, (double *)c->psi + 2 * c->ps * p0, buf.data(), 2 * sizeof(double), (int)np, false,
                                 hipMemcpyDeviceToHost));
            HIPCHK(c, hipStreamSynchronize(c->stream));
        } else {
            HIPCHK(c, plane_copy(c, (float *)c->psi + 2 * c->ps * p0, fbuf.data(), 2 * sizeof(float), (int)np, false,
                                 hipMemcpyDeviceToHost));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            for (size_t i = 0; i < cnt; ++i) buf[i] = std::ldexp((double)fbuf[i], c->psi_scale);
        }
        if (std::fwrite(buf.data(), sizeof(double), cnt, f.f) != cnt) {
            c->err = std::string("write failed: ") + path;
            return IFDM_E_IO;
        }
    }
    return IFDM_OK;
}

// read_rerun (src/gadget_IC.h:709-750): a, x and psi from output<t>.dat
int ifdm_read_snapshot(ifdm_ctx *c, const char *path, double *a_out, double *mfdm_out, double *g_out) {
    if (!c || !path) return IFDM_E_ARG;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    File f(path, "rb");
    if (!f.f) { c->err = std::string("cannot find ") + path; return IFDM_E_IO; }
    const int n = c->n;
    double hdr[3];
    std::vector<double> x(n);
    if (std::fread(hdr, 8, 3, f.f) != 3 || std::fread(x.data(), 8, n, f.f) != (size_t)n) {
        c->err = std::string("truncated snapshot header: ") + path;
        return IFDM_E_FORMAT;
    }
    const size_t plane = (size_t)n * n;
    const size_t planes_per = std::max<size_t>(1, (size_t(1) << 27) / plane);
    std::vector<double> buf(2 * plane * planes_per);
    std::vector<float> fbuf(c->mixed ? 2 * plane * planes_per : 0);
    double norm = 0.0;
    for (size_t p0 = 0; p0 < (size_t)n; p0 += planes_per) {
        const size_t np = std::min(planes_per, (size_t)n - p0), cnt = 2 * plane * np;
        if (std::fread(buf.data(), sizeof(double), cnt, f.f) != cnt) {
            c->err = std::string("truncated snapshot: ") + path;
            return IFDM_E_FORMAT;
        }
        if (!c->mixed) {
            HIPCHK(c, plane_copy(c, (double *)c->psi + 2 * c->ps * p0, buf.data(), 2 * sizeof(double), (int)np, true,
                                 hipMemcpyHostToDevice));
            HIPCHK(c, hipStreamSynchronize(c->stream));
        } else {
            for (size_t i = 0; i < cnt; ++i) {
                fbuf[i] = (float)std::ldexp(buf[i], -c->psi_scale);
                norm += buf[i] * buf[i];
            }
            HIPCHK(c, plane_copy(c, (float *)c->psi + 2 * c->ps * p0, fbuf.data(), 2 * sizeof(float), (int)np, true,
                                 hipMemcpyHostToDevice));
            HIPCHK(c, hipStreamSynchronize(c->stream));
        }
    }
    c->pending = false;
    c->ck_nsteps = -1;   // psi replaced: the adjoint checkpoints no longer apply
    if (c->mixed) c->rho_mean = c->ms * norm / (double)c->n3;
    if (a_out) *a_out = hdr[2];
    if (mfdm_out) *mfdm_out = hdr[0];
    if (g_out) *g_out = hdr[1];
    return IFDM_OK;
}

}  // extern "C"
