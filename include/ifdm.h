/*
 * ifdm.h -- C ABI of libifdm.so, the MI355X (gfx950) Schrodinger-Poisson
 * kick-drift-kick solver.
 *
 * The reference (jchan192/IFDM_Code) has no plugin or FFI layer: its hot path
 * is a set of C++ free functions over FFTW-MPI plans.  Each entry point below
 * names the reference function it replaces.  Conventions:
 *   - every function returns 0 on success or a negative IFDM_E* code; the
 *     message is available from ifdm_last_error(ctx).  Nothing calls exit().
 *   - host buffers are caller-owned and copied; the context owns all device
 *     memory and one HIP stream; calls are synchronous with respect to host
 *     buffers.  A context is not thread-safe.
 *   - array layout is the reference's: C order (x, y, z), z fastest,
 *     offset k + N*(j + N*i) (src/CosmoBEC_MPI_stan.cpp:133); psi is
 *     interleaved complex (re, im) doubles; phi is returned unpadded (the
 *     reference's padded row of 2*(N/2+1) doubles minus its 2 pad entries,
 *     src/CosmoBEC_MPI_stan.cpp:113-118).
 */
#ifndef IFDM_H
#define IFDM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IFDM_ABI_VERSION 1

enum {
    IFDM_OK = 0,
    IFDM_E_ARG = -1,      /* invalid argument (size, null pointer, mode) */
    IFDM_E_HIP = -2,      /* HIP runtime error */
    IFDM_E_NOMEM = -3,    /* device or host allocation failed */
    IFDM_E_IO = -4,       /* file open / read / write failed */
    IFDM_E_FORMAT = -5,   /* malformed input file */
    IFDM_E_STATE = -6     /* call not valid in the current state */
};

enum { IFDM_FP64 = 0, IFDM_MIXED = 1 };

/* Physical and numerical configuration of one context.  Units are the
 * reference's internal ones after its argv conversions
 * (src/CosmoBEC_MPI_stan.cpp:329-332, src/Cosmoinput3D.h:2-17). */
typedef struct {
    int n;                 /* mesh points per axis (res); power of two, 8..1024 */
    int precision;         /* IFDM_FP64 or IFDM_MIXED */
    double xmax;           /* half box length in Mpc (argv xmax / h_0) */
    double mfdm;           /* particle mass in eV */
    double g;              /* self-interaction in eV Mpc^3 (argv g / cm2Mpc^3) */
    double rhoc;           /* mean density subtracted in the Poisson source, Msun/Mpc^3 */
    int ref_maxphi_range;  /* 1: max|phi| over the first N^3 entries of the
                              reference's padded phi array (its one-rank
                              std::max_element range); 0: all cells */
    int device;            /* HIP device ordinal */
} ifdm_config;

/* Reductions after a Poisson solve / step: the two scans of
 * Timestep_criteria (src/CosmoBEC_MPI_stan.cpp:26-27) plus the norm. */
typedef struct {
    double max_phi;        /* max |phi| */
    double max_rho;        /* max |psi|^2 */
    double norm;           /* sum |psi|^2 (FP64 accumulation in both modes) */
} ifdm_stats;

/* Time integrals of one step (par.dt_KE, par.dt_PE, par.dt_QI,
 * src/CosmoBEC_MPI_stan.cpp:66-68). */
typedef struct {
    double dKE, dPE, dQI;
} ifdm_dt;

typedef struct ifdm_ctx ifdm_ctx;

/* Replaces the Operators constructor (src/struct3D.h:105-124) and Pre_FFTW
 * (src/CosmoBEC_MPI_stan.cpp:160-182): allocates psi, phi and the
 * half-spectrum on the device and builds the twiddle tables. */
int ifdm_create(const ifdm_config *cfg, ifdm_ctx **out);
void ifdm_destroy(ifdm_ctx *ctx);
const char *ifdm_last_error(const ifdm_ctx *ctx);
/* Message of the last ifdm_create failure (no context exists then). */
const char *ifdm_create_error(void);

/* psi in / out, host complex128 buffers of N^3 (re, im) pairs in the
 * reference layout (the opr.wfc vector, src/struct3D.h:102). */
int ifdm_set_psi(ifdm_ctx *ctx, const double *psi_c128);
int ifdm_get_psi(ifdm_ctx *ctx, double *psi_c128);
/* phi out, N^3 doubles (opr.phi with its pad columns dropped). */
int ifdm_get_phi(ifdm_ctx *ctx, double *phi_f64);
/* phi in (N^3 doubles), e.g. for a restart that carries phi. */
int ifdm_set_phi(ifdm_ctx *ctx, const double *phi_f64);

/* poisson(), src/CosmoBEC_MPI_stan.cpp:80-119: phi from |psi|^2. */
int ifdm_poisson(ifdm_ctx *ctx, ifdm_stats *stats);

/* One half-kick with the current phi, src/CosmoBEC_MPI_stan.cpp:130-134:
 * psi *= exp(-i/2 (dPE m/(c^2 hbar) phi + dQI g/hbar |psi|^2)). */
int ifdm_kick(ifdm_ctx *ctx, double dPE, double dQI);

/* Split_op_MPI, src/CosmoBEC_MPI_stan.cpp:121-159: one full step
 * K(1/2) - drift - poisson - K(1/2) with the given integrals.  phi must be
 * current on entry (ifdm_poisson).  stats: reductions of the post-step phi
 * and psi (what the next Timestep_criteria reads); may be NULL. */
int ifdm_kdk_step(ifdm_ctx *ctx, const ifdm_dt *dt, ifdm_stats *stats);

/* nsteps consecutive Split_op_MPI steps with a given schedule (replay of a
 * recorded run), merging the second half-kick of step n with the first of
 * step n+1 into one pass.  The state on return equals nsteps calls of
 * ifdm_kdk_step.  stats_per_step (nsteps entries) may be NULL. */
int ifdm_run(ifdm_ctx *ctx, const ifdm_dt *sched, int nsteps, ifdm_stats *stats_per_step);

/* The reference's time loop (Simulation, src/CosmoBEC_MPI_stan.cpp:212-304)
 * without its snapshot triggers: a Poisson solve, then per step
 * Timestep_criteria (host, FP64) and Split_op_MPI, with the second half-kick
 * of each step merged into the first of the next.  a0: scale factor at the
 * start; a_out (may be NULL): scale factor after the last step; sched_out
 * (nsteps entries, may be NULL): the integrals used, for replay. */
int ifdm_simulate(ifdm_ctx *ctx, double a0, int nsteps, double *a_out, ifdm_dt *sched_out);

/* Times each pass of one step with HIP events on the context's stream,
 * averaged over iters repetitions (ms_out[IFDM_NPASS]; passes A..I of the
 * fused schedule, then the final reduction).  The state is advanced. */
#define IFDM_NPASS 10
int ifdm_time_passes(ifdm_ctx *ctx, const ifdm_dt *dt, int iters, float *ms_out);

/* Live per-pass timing: with on != 0 every step run by ifdm_run or
 * ifdm_simulate brackets each pass with HIP events on the context's stream
 * and accumulates the durations (reset by each call of this function).
 * ifdm_get_profile returns the summed ms per pass and the number of steps. */
int ifdm_set_profiling(ifdm_ctx *ctx, int on);
int ifdm_get_profile(const ifdm_ctx *ctx, double *ms_sum, long *count);

/* Streaming-bandwidth reference: copies `bytes` between two scratch device
 * buffers `iters` times with a 16-byte-vector kernel and reports the
 * achieved (read + write) GB/s.  Allocates and frees the two buffers. */
int ifdm_stream_copy(ifdm_ctx *ctx, size_t bytes, int iters, double *gbs);

/* The drift alone (forward 3-D FFT, exp(-i KE_phase dKE)/N^3, inverse),
 * src/CosmoBEC_MPI_stan.cpp:136-144. */
int ifdm_drift(ifdm_ctx *ctx, double dKE);

/* Unnormalised in-place 3-D c2c transform of psi, sign -1 (FFTW_FORWARD)
 * or +1 (FFTW_BACKWARD): the wfc_fplan / wfc_bplan executes of
 * src/CosmoBEC_MPI_stan.cpp:136,144. */
int ifdm_fft3d(ifdm_ctx *ctx, int sign);

/* ---- host-side physics (no device work) ---------------------------- */

/* Timestep_criteria, src/CosmoBEC_MPI_stan.cpp:19-78: from the reductions
 * and the current scale factor a, the three time integrals (adaptive
 * Gauss-Kronrod 21, relative tolerance 1e-10, as gsl_integration_qags) and
 * a_next.  which_out (may be NULL) receives 3, 2 or 4 as the reference's n. */
int ifdm_timestep_criteria(const ifdm_config *cfg, double a, const ifdm_stats *stats,
                           ifdm_dt *dt_out, double *a_next, int *which_out);

/* integral of 1/(H(x) x^power) dx from a0 to a1 (Ha{2,3,4}_preintegrat,
 * src/struct3D.h:127-146) by the same adaptive Gauss-Kronrod rule. */
double ifdm_scale_integral(double a0, double a1, int power);

/* Constants of src/Cosmoinput3D.h:2-17 by name ("hbar", "c", "bigG",
 * "H_0", "h_0", "Omega_m0", "rhoc", "cm2Mpc", "kg2eV", "kg2solar",
 * "s2Myr", "m2Mpc"); NaN for an unknown name. */
double ifdm_constant(const char *name);

int ifdm_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif
