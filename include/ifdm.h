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

#define IFDM_ABI_VERSION 2

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
    int plane_pad;         /* device layout of psi and phi: cells of padding
                              after each x-plane of N*N cells, rounded up to
                              a multiple of 8 (0: the built-in 256 bytes,
                              16 cells in FP64 and 32 in mixed;
                              negative: none; at most 65536).  Host arrays
                              stay dense. */
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
 * reference layout (the opr.wfc vector, src/struct3D.h:102).  In mixed
 * mode the precision conversion runs on up to 16 host threads that the
 * call starts and joins before it returns (also in ifdm_set_phi,
 * ifdm_get_phi and ifdm_adjoint). */
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

/* One Split_op_MPI step with the second half-kick deferred: the first
 * half-kick of this step is merged with the deferred half-kick of the
 * previous ifdm_step (one phase multiply), then drift and Poisson; this
 * step's second half-kick stays pending.  Every call that observes or
 * changes psi from outside (ifdm_get_psi, ifdm_write_snapshot, ifdm_kick,
 * ifdm_drift, ifdm_fft3d, ifdm_kdk_step, ifdm_set_phi) applies a pending
 * half-kick first (ifdm_flush); ifdm_set_psi discards it.  stats: the
 * post-step reductions (unchanged by the pending kick); may be NULL. */
int ifdm_step(ifdm_ctx *ctx, const ifdm_dt *dt, ifdm_stats *stats);
int ifdm_flush(ifdm_ctx *ctx);

/* nsteps consecutive Split_op_MPI steps with a given schedule (replay of a
 * recorded run), merging the second half-kick of step n with the first of
 * step n+1 into one pass.  The state on return equals nsteps calls of
 * ifdm_kdk_step.  stats_per_step (nsteps entries) may be NULL. */
int ifdm_run(ifdm_ctx *ctx, const ifdm_dt *sched, int nsteps, ifdm_stats *stats_per_step);

/* The reference's time loop (Simulation, src/CosmoBEC_MPI_stan.cpp:212-304)
 * without its snapshot triggers: a Poisson solve, then per step
 * Timestep_criteria (FP64) and Split_op_MPI, with the second half-kick of
 * each step merged into the first of the next.  a0: scale factor at the
 * start; a_out (may be NULL): scale factor after the last step; sched_out
 * (nsteps entries, may be NULL): the integrals used, for replay.  By default
 * (option "device_loop" 1) the time step is decided on the device between
 * the steps, with the arithmetic of ifdm_timestep_criteria, so the host
 * issues every step without waiting; the schedule and the states equal the
 * host-driven loop (device_loop 0: ifdm_timestep_criteria and ifdm_step
 * per step) bit for bit. */
int ifdm_simulate(ifdm_ctx *ctx, double a0, int nsteps, double *a_out, ifdm_dt *sched_out);

/* ifdm_simulate with flags, so one run can be driven in pieces (to time a
 * stretch of it, or to return to the caller between steps without a
 * flush): IFDM_SIM_KEEP_PENDING leaves the last step's second half-kick
 * pending (no ifdm_flush at the end); IFDM_SIM_RESUME continues such a run
 * without the entry Poisson solve, taking phi and the reductions from its
 * last step (IFDM_E_STATE when no half-kick is pending).  Calls with
 * KEEP_PENDING then RESUME give the same states as one call over all the
 * steps.  flags = 0 is ifdm_simulate. */
#define IFDM_SIM_RESUME 1
#define IFDM_SIM_KEEP_PENDING 2
int ifdm_simulate_ex(ifdm_ctx *ctx, double a0, int nsteps, double *a_out, ifdm_dt *sched_out, int flags);

/* Times each pass of one step with HIP events on the context's stream,
 * averaged over iters repetitions (ms_out[IFDM_NPASS]: the drift table (in
 * the on-device loop of ifdm_simulate, the step preparation: the final
 * reduction, the time step and the drift table); the
 * passes A, B, C - the y, x, z drift round trips, A with the kick, C with
 * the Poisson source and its z transform; F, G, H, I - the Poisson solve on
 * the half spectrum; then the final reduction).  The state is advanced. */
#define IFDM_NPASS 9
int ifdm_time_passes(ifdm_ctx *ctx, const ifdm_dt *dt, int iters, float *ms_out);

/* Options.
 *   "ic_deposit": 0 (default) deposits the initial conditions the way the
 *   reference's CIC_scheme does at one rank, bit for bit: each cell's
 *   contributions in the reference's particle order, the density in FP64
 *   and the three velocity grids in float (src/gadget_IC.h:112-150,583),
 *   the divergence in float (:170), so two runs give the same psi; 1 uses
 *   unordered FP64 atomic adds into FP64 grids (round-off differences from
 *   run to run).  Set before ifdm_ic_begin / ifdm_read_gadget.
 * Launch-schedule options (no effect on results beyond round-off):
 *   "persistent" mask: bit 0 the y passes on psi (drift only; the kick pass
 *   always runs one tile per work-group), bit 1 the x passes on psi, bit 2
 *   the forward y pass on the half spectrum (F), bit 3 its x pass (G), bit 4
 *   its inverse y pass (H); a set bit launches a resident set of
 *   work-groups that walk the tiles and prefetch the next tile's data during
 *   the current transform instead of one work-group per tile.  Default: 0
 *   in both precisions.
 *   "device_loop": 1 (default) ifdm_simulate / ifdm_simulate_ex decide the
 *   time step on the device; 0 on the host after each step (the same
 *   results; a host wait per step). */
int ifdm_set_option(ifdm_ctx *ctx, const char *name, long value);

/* Live per-pass timing: with on != 0 every step run by ifdm_run or
 * ifdm_simulate brackets each pass with HIP events on the context's stream
 * and accumulates the durations (reset by each call of this function).
 * ifdm_get_profile returns the summed ms per pass and the number of steps.
 * Option "profile_mask" (ifdm_set_option; bit p = pass p of ifdm_time_passes'
 * order, default all) limits the brackets to the chosen passes, so a timed
 * run carries two events per step around one pass only. */
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

/* ---- diagnostics (no counterpart in the reference) ------------------- */

/* Matter power spectrum of delta = m_s|psi|^2 / rho_mean - 1 on shells of
 * width 2 pi / L: bin b (0-based) holds the modes with round(|n|) = b + 1,
 * n the integer wave vector.  pk[b] = L^3 <|delta_k|^2> / N^6 (Mpc^3),
 * k_mean[b] the mode-weighted mean |k| (1/Mpc), nmodes[b] the number of
 * modes (both halves of the Hermitian spectrum counted).  k_mean and
 * nmodes may be NULL.  Uses the Poisson scratch; psi and phi are kept. */
int ifdm_power_spectrum(ifdm_ctx *ctx, int nbins, double *k_mean, double *pk, double *nmodes);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (tight diagnostic loops, code that must not call hipMalloc): the
 * same P(k) as ifdm_power_spectrum, with the same outputs, but it allocates
 * nothing.  Its bins live in a fixed scratch array of the library on each
 * GPU (6 x 4096 doubles, 192 KiB, held until the process ends), cleared by
 * a kernel on every call; nbins is at most 4096 (IFDM_E_ARG beyond).  Every
 * call on one GPU shares that array: calls that overlap on one GPU (two
 * contexts on the same device, two threads) race. */
int ifdm_power_spectrum_noalloc(ifdm_ctx *ctx, int nbins, double *k_mean, double *pk, double *nmodes);

/* ---- adjoint (no counterpart in the reference) ----------------------- */

/* Forward run of nsteps steps with the given schedule (as ifdm_run) that
 * keeps psi every K steps (step 0, K, 2K, ...) on the device for
 * ifdm_adjoint.  phi is solved for the current psi first, so the run is a
 * function of psi alone.  In mixed contexts the checkpoints are FP32 and
 * ifdm_adjoint holds g as 2^-s g in FP32 (the sweep is linear in g).  The
 * checkpoints serve one reverse sweep; any call that changes psi (set,
 * step, kick, drift, transform, IC, snapshot read) discards them. */
int ifdm_run_checkpointed(ifdm_ctx *ctx, const ifdm_dt *sched, int nsteps, int K);

/* Reverse sweep of the last ifdm_run_checkpointed (the same sched, checked
 * entry by entry; IFDM_E_STATE otherwise or when the checkpoints were used
 * or discarded).
 * On entry g holds dL/dRe psi_T + i dL/dIm psi_T for a real loss L of the
 * final state (host N^3 complex128, reference layout); on return it holds
 * dL/dRe psi_0 + i dL/dIm psi_0.  The schedule is frozen (dt does not
 * depend on psi here).  Each kick psi' = psi e^{-i theta}, theta =
 * aphi phi + arho |psi|^2, is reversed exactly (it preserves |psi|) with
 * g = e^{i theta} g' + 2 psi (aphi G(m_s s) + arho s), s = Im(conj(g') psi'),
 * G the Poisson Green's operator; each drift by its conjugate; psi is
 * reverse-evolved and replaced by the stored checkpoint every K steps.
 * psi holds psi_0 on return. */
int ifdm_adjoint(ifdm_ctx *ctx, const ifdm_dt *sched, int nsteps, double *g_c128);
/* The same with g in device memory (complex128 N^3, e.g. a torch tensor);
 * FP64 contexts. */
int ifdm_adjoint_device(ifdm_ctx *ctx, const ifdm_dt *sched, int nsteps, void *g_dev);

/* psi from / to device memory (complex128 N^3 on the context's device). */
int ifdm_set_psi_device(ifdm_ctx *ctx, const void *psi_dev);
int ifdm_get_psi_device(ifdm_ctx *ctx, void *psi_dev);

/* ---- initial conditions and snapshots -------------------------------- */

/* What the Gadget-1 IC header gave (src/gadget_IC.h:289-329). */
typedef struct {
    double box;            /* BoxSize, kpc/h */
    double pmass;          /* mass[1] * 1e10, Msun/h (:317) */
    double a;              /* time: the scale factor of the IC (:318) */
    double rhoc;           /* mean density derived from the header (:319), Msun/Mpc^3 */
    long long npart;       /* particles read */
    long long skipped;     /* particles outside [-Lg/2, box) in some coordinate, not deposited */
} ifdm_ic_info;

/* read_gadget, src/gadget_IC.h:573-672 at one rank: reads the nchunks
 * files <icpath>.0 .. (4 in the reference, :610), checks each BoxSize
 * against 2 xmax 1e3 h_0 (:309; IFDM_E_FORMAT instead of exit), deposits
 * every particle with the 8-cell CIC on the device (CIC_scheme, :45-156),
 * forms the mass-weighted velocities, the velocity-divergence phase
 * (initial_velocity_Poisson_Solver, :157-247) and psi = 1.225 sqrt(rho/m_s)
 * exp(i m theta / hbar) (:666), and sets the context's rho_c (par.rhoc).
 * ref_ghosts = 1 also deposits the trailing particles of each x-sorted chunk
 * beyond the last cell centre a second time, as the reference does at one
 * rank (:445-448).  voutput_path (may be NULL): append the float vx grid
 * there (write_binary_vg, src/write3D.h:24-53). */
int ifdm_read_gadget(ifdm_ctx *ctx, const char *icpath, int nchunks, int ref_ghosts,
                     const char *voutput_path, ifdm_ic_info *info);

/* The same IC pipeline from particles already in host memory:
 * ifdm_ic_begin (box kpc/h, particle mass Msun/h), any number of
 * ifdm_ic_deposit calls (np particles, pos and vel as 3*np floats in kpc/h
 * and Gadget velocity units), then ifdm_ic_finish with the scale factor.
 * theta_out (N^3 complex, may be NULL) receives the phase field opr.phik,
 * vx_out (N^3 floats, may be NULL) the mass-weighted vx grid. */
int ifdm_ic_begin(ifdm_ctx *ctx, double box_kpc_h, double pmass_msun_h);
int ifdm_ic_deposit(ifdm_ctx *ctx, const float *pos, const float *vel, long long np, int ref_ghosts);
int ifdm_ic_finish(ifdm_ctx *ctx, double a, double *theta_out, float *vx_out, long long *skipped_out);

/* rho_c subtracted in the Poisson source (par.rhoc; the rerun path sets
 * the constant 3.51059e10, src/gadget_IC.h:744). */
int ifdm_set_rhoc(ifdm_ctx *ctx, double rhoc);

/* writeall (src/write3D.h:86-97): output file [f64 mfdm eV][f64 g eV cm^3]
 * [f64 a][f64 x[N]][complex128 psi N^3]; psi is streamed plane by plane. */
int ifdm_write_snapshot(ifdm_ctx *ctx, const char *path, double a, double mfdm_ev, double g_ev_cm3);
/* read_rerun (src/gadget_IC.h:709-750): psi and a from such a file. */
int ifdm_read_snapshot(ifdm_ctx *ctx, const char *path, double *a_out, double *mfdm_out, double *g_out);

/* ---- host-side physics (no device work) ---------------------------- */

/* Timestep_criteria, src/CosmoBEC_MPI_stan.cpp:19-78: from the reductions
 * and the current scale factor a, the three time integrals (adaptive
 * Gauss-Kronrod 21, relative tolerance 1e-10, as gsl_integration_qags) and
 * a_next.  which_out (may be NULL) receives 3, 2 or 4 as the reference's n. */
int ifdm_timestep_criteria(const ifdm_config *cfg, double a, const ifdm_stats *stats,
                           ifdm_dt *dt_out, double *a_next, int *which_out);

/* Snapshot triggers of the time loop (Simulation,
 * src/CosmoBEC_MPI_stan.cpp:238-284; variant 1 = src/CosmoBEC_MPI.cpp,
 * which differs at :250 and :256-261).  Called once per step after
 * Timestep_criteria, with the updated a.  Returns 1 when a snapshot is due
 * (its number in *t_out), 0 otherwise, -1 on bad arguments; *stop = 1 when
 * the loop ends after this write (the reference's `break`).  a_triggers = 0
 * keeps only the every-twrite / last-step trigger (:282).  The state starts
 * as {0, 0, 2} (write_num, sten_counter, write_sten_counter, :208-210). */
typedef struct {
    int write_num;
    int sten_counter;
    int write_sten_counter;
} ifdm_trigger_state;
int ifdm_output_trigger(int variant, int a_triggers, double a, int tn, int twrite, int t_stop,
                        ifdm_trigger_state *st, int *t_out, int *stop);

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
