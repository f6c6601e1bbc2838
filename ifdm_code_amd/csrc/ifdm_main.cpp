// ifdm_main.cpp -- command-line driver with the reference's ten run-time
// parameters (src/CosmoBEC_MPI_stan.cpp:313-379, README.md:8-20):
//
//   ifdm mfdm g xmax res t_write t_stop ti opath icpath rerun [options]
//
// mfdm [eV], g [eV cm^3], xmax = box/2 [Mpc/h], res (mesh points per axis),
// t_write, t_stop, ti (first step; the snapshot number to restart from when
// rerun = 1), opath (output directory prefix), icpath (Gadget IC prefix, or
// the directory of the restart snapshot), rerun (0/1).  Options:
//   --precision fp64|mixed   storage / FFT precision (default fp64)
//   --no-a-triggers          only the every-t_write / last-step snapshots
//   --variant stan|mpi       trigger chain of CosmoBEC_MPI_stan.cpp (default) or CosmoBEC_MPI.cpp
//   --ref-maxphi-range       max|phi| over the reference's one-rank scan range
//   --ref-ghosts             deposit the x-sorted chunks' trailing particles twice, as at one rank
//   --nchunks K              IC files per snapshot (default 4)
//   --ic-deposit reference|atomic
//                            CIC in the reference's particle order with float
//                            velocity grids (default; the reference's bits) or
//                            unordered FP64 atomics
//   --device D               HIP device
// Uses only the C ABI of libifdm.so.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>

#include "../../include/ifdm.h"

static void usage() {
    std::fprintf(stderr,
                 "usage: ifdm mfdm g xmax res t_write t_stop ti opath icpath rerun "
                 "[--precision fp64|mixed] [--no-a-triggers] [--variant stan|mpi] [--ref-maxphi-range] "
                 "[--ref-ghosts] [--nchunks K] [--ic-deposit reference|atomic] [--device D]\n");
}

static int check(int rc, ifdm_ctx *ctx, const char *what) {
    if (rc) std::fprintf(stderr, "ERROR: %s failed (%d): %s\n", what, rc, ctx ? ifdm_last_error(ctx) : ifdm_create_error());
    return rc;
}

int main(int argc, char **argv) {
    if (argc < 11) {
        std::fprintf(stderr, "ERROR: Arguments %d != 11\n", argc);
        usage();
        return 2;
    }
    const double mfdm = std::atof(argv[1]);
    const double g_in = std::atof(argv[2]);
    const double xmax_in = std::atof(argv[3]);
    const int res = std::atoi(argv[4]);
    const int twrite = (int)std::atof(argv[5]);
    const int t_stop = (int)std::atof(argv[6]);
    const int ti = (int)std::atof(argv[7]);
    const std::string opath = argv[8], icpath = argv[9];
    const bool rerun = std::atof(argv[10]) == 1;
    int precision = IFDM_FP64, a_triggers = 1, variant = 0, ref_range = 0, ref_ghosts = 0, nchunks = 4, device = 0;
    int ic_atomic = 0;
    for (int i = 11; i < argc; ++i) {
        const std::string o = argv[i];
        auto next = [&](void) -> const char * {
            if (i + 1 >= argc) { usage(); std::exit(2); }
            return argv[++i];
        };
        if (o == "--precision") {
            const std::string v = next();
            if (v == "fp64") precision = IFDM_FP64;
            else if (v == "mixed") precision = IFDM_MIXED;
            else { std::fprintf(stderr, "ERROR: unknown precision %s\n", v.c_str()); usage(); return 2; }
        } else if (o == "--no-a-triggers") a_triggers = 0;
        else if (o == "--variant") {
            const std::string v = next();
            if (v == "stan") variant = 0;
            else if (v == "mpi") variant = 1;
            else { std::fprintf(stderr, "ERROR: unknown variant %s\n", v.c_str()); usage(); return 2; }
        }
        else if (o == "--ref-maxphi-range") ref_range = 1;
        else if (o == "--ref-ghosts") ref_ghosts = 1;
        else if (o == "--nchunks") nchunks = std::atoi(next());
        else if (o == "--ic-deposit") {
            const std::string v = next();
            if (v == "reference") ic_atomic = 0;
            else if (v == "atomic") ic_atomic = 1;
            else { std::fprintf(stderr, "ERROR: unknown deposit %s\n", v.c_str()); usage(); return 2; }
        }
        else if (o == "--device") device = std::atoi(next());
        else { usage(); return 2; }
    }
    // argv unit conversions, src/CosmoBEC_MPI_stan.cpp:329-332
    ifdm_config cfg;
    std::memset(&cfg, 0, sizeof(cfg));
    cfg.n = res;
    cfg.precision = precision;
    cfg.xmax = xmax_in / ifdm_constant("h_0");
    cfg.mfdm = mfdm;
    cfg.g = g_in / std::pow(ifdm_constant("cm2Mpc"), 3);
    cfg.rhoc = ifdm_constant("rhoc");
    cfg.ref_maxphi_range = ref_range;
    cfg.device = device;
    std::printf(" Input parameters: \nmfdm=%g\ng=%g\nxmax = %g Mpc \nBoxsize = %g Mpc \nres = %d\ntwrite = %d\n"
                "timestep = %d\nti = %d\nopath = %s\n\n",
                mfdm, g_in, cfg.xmax, 2 * cfg.xmax, res, twrite, t_stop, ti, opath.c_str());
    ifdm_ctx *ctx = nullptr;
    if (check(ifdm_create(&cfg, &ctx), nullptr, "ifdm_create")) return 1;
    if (check(ifdm_set_option(ctx, "ic_deposit", ic_atomic), ctx, "ifdm_set_option")) { ifdm_destroy(ctx); return 1; }
    const time_t t0 = time(nullptr);
    double a = 0.0;
    int rc = 0;
    if (!rerun) {
        ifdm_ic_info info;
        const std::string vout = opath + "voutput.dat";
        rc = ifdm_read_gadget(ctx, icpath.c_str(), nchunks, ref_ghosts, vout.c_str(), &info);
        if (check(rc, ctx, "ifdm_read_gadget")) { ifdm_destroy(ctx); return 1; }
        a = info.a;
        std::printf("npart = %lld\nmass = %g\na = %g\nBoxSize = %g\npar.rhoc = %g M_solar / Mpc**3\n", info.npart,
                    info.pmass, info.a, info.box, info.rhoc);
        if (info.skipped) std::printf("WARNING: %lld particles outside the box were not deposited\n", info.skipped);
    } else {
        const std::string snap = icpath + "output" + std::to_string(ti) + ".dat";   // src/gadget_IC.h:716
        rc = ifdm_read_snapshot(ctx, snap.c_str(), &a, nullptr, nullptr);
        if (check(rc, ctx, "ifdm_read_snapshot")) { ifdm_destroy(ctx); return 1; }
        ifdm_set_rhoc(ctx, ifdm_constant("rhoc"));                                    // src/gadget_IC.h:744
    }
    ifdm_stats st;
    if (check(ifdm_poisson(ctx, &st), ctx, "ifdm_poisson")) { ifdm_destroy(ctx); return 1; }
    ifdm_trigger_state trig = {0, 0, 2};
    for (int tn = ti; tn < t_stop; ++tn) {
        const time_t ts = time(nullptr);
        ifdm_dt dt;
        double a_next;
        int which = 0;
        if ((rc = check(ifdm_timestep_criteria(&cfg, a, &st, &dt, &a_next, &which), ctx, "ifdm_timestep_criteria")))
            break;
        std::printf("ai = %g a_next = %g da = %g n = %d ", a, a_next, a_next - a, which);
        a = a_next;                                                                   // :72
        int tw = -1, stop = 0;
        if (ifdm_output_trigger(variant, a_triggers, a, tn, twrite, t_stop, &trig, &tw, &stop) == 1) {
            const std::string out = opath + "output" + std::to_string(tw) + ".dat";
            if ((rc = check(ifdm_write_snapshot(ctx, out.c_str(), a, mfdm, g_in), ctx, "ifdm_write_snapshot"))) break;
            std::printf("write %s at step %d ", out.c_str(), tn + 1);
            if (stop) { std::printf("\n"); break; }
        }
        std::printf("tn = %d\n", tn);
        if ((rc = check(ifdm_step(ctx, &dt, &st), ctx, "ifdm_step"))) break;
        std::printf("simulation takes:%ld seconds\n", (long)(time(nullptr) - ts));
    }
    if (!rc) rc = check(ifdm_flush(ctx), ctx, "ifdm_flush");
    std::printf("simulation takes:%ld seconds\n", (long)(time(nullptr) - t0));
    ifdm_destroy(ctx);
    return rc ? 1 : 0;
}
