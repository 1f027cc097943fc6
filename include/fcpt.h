/*
 * fcpt.h -- C ABI of the MI355X (gfx950) hydro-step library libfcpt_hip.so.
 *
 * One context (fcpt_ctx) holds one 2-D polar disk resident in HBM: the evolved
 * fields (Sigma, v_r, v_phi, E), the fields derived from them, the damping
 * reference fields and the per-ring metric arrays.  Every entry point below
 * replaces one function of the reference's per-timestep hydro update; the
 * reference function and its declaration (file:line in the reference tree)
 * are named in the comment above each declaration.
 *
 * Layout contract (identical to the reference's t_polargrid, polargrid.h:111-123):
 *   scalar fields   [nr][nphi]      row-major, phi contiguous
 *   v_r, tau_r_phi  [nr + 1][nphi]  (the reference's "vector" grids)
 *   ring 0 and ring nr-1 are ghost rings; Radii[1] = Rmin, Radii[nr-1] = Rmax.
 * Host arrays passed to upload/download use exactly this layout.
 *
 * Errors: every function returns 0 on success or a negative fcpt_status; the
 * message of the last error on a context is returned by fcpt_last_error().
 * No entry point aborts the process.  A context may be used by one host
 * thread at a time; distinct contexts may be used concurrently.
 */
#ifndef FCPT_H
#define FCPT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fcpt_ctx fcpt_ctx;

typedef enum fcpt_status {
    FCPT_OK = 0,
    FCPT_ERR_ARG = -1,      /* invalid argument or size */
    FCPT_ERR_DEVICE = -2,   /* HIP runtime error */
    FCPT_ERR_STATE = -3,    /* call order / missing state */
    FCPT_ERR_NONFINITE = -4 /* CFL reduction met a NaN or non-positive dt; reported by the
                               first call that reads it back, which clears it (as do
                               fcpt_set_time, fcpt_upload_state and fcpt_read_snapshot) */
} fcpt_status;

/* artificial viscosity (reference parameters.h:156-161) */
enum { FCPT_ARTVISC_NONE = 0, FCPT_ARTVISC_TW = 1, FCPT_ARTVISC_SN = 2 };
/* composite inner/outer boundary (reference boundary_conditions/config.cpp) */
enum { FCPT_BC_REFLECTING = 0, FCPT_BC_ZEROGRADIENT = 1, FCPT_BC_OUTFLOW = 2 };
/* beta-cooling reference energy (reference parameters.cpp:443-457) */
enum { FCPT_BETA_REF_ZERO = 0, FCPT_BETA_REF_REFERENCE = 1, FCPT_BETA_REF_FLOOR = 3 };
/* storage of the 2-D fields.  Arithmetic is FP64 in both modes.
   FCPT_PREC_FP64:    every field in FP64 (results match the reference's host build).
   FCPT_PREC_SPLIT32: every 2-D field but the gas potential stored as an FP32
                      deviation from an FP64 per-ring background (value =
                      background[ring] + deviation), halving the HBM traffic of
                      the step; the potential stays FP64.  The backgrounds are
                      the ring means of the state given to fcpt_upload_state /
                      fcpt_upload_field and the profiles derived from them;
                      downloads return background + deviation in FP64.  Stored
                      deviations are rounded stochastically by default
                      (FCPT_PARAM_SPLIT_ROUNDING). */
enum { FCPT_PREC_FP64 = 0, FCPT_PREC_SPLIT32 = 1 };
/* hydro integrator (reference types.h:64-65, Interpret.cpp:173-186): the
   LeapFrog scheme also scales the artificial-viscosity, viscosity and
   heating/cooling limits of the CFL condition by 0.6 (cfl.cpp:254-261) */
enum { FCPT_INTEGRATOR_EULER = 0, FCPT_INTEGRATOR_LEAPFROG = 1 };
/* potential smoothing switches of compute_smoothing (Force.cpp:145-159):
   CompatibilityNoStarSmoothing (body 0 unsmoothed) and
   CompatibilitySmoothingPlanetLoc (each body's smoothing fixed at
   ThicknessSmoothing * AspectRatio * r_k^(1 + FlaringIndex), r_k = hypot(x_k, y_k),
   instead of ThicknessSmoothing * H of the cell); the same bit values as the
   disk-force flags FCPT_FORCE_NO_STAR_SMOOTHING / FCPT_FORCE_SMOOTHING_PLANETLOC */
enum { FCPT_SMOOTHING_NO_STAR = 1, FCPT_SMOOTHING_PLANETLOC = 2 };

/* field identifiers for fcpt_upload_field / fcpt_download_field */
typedef enum fcpt_field {
    FCPT_SIGMA = 0, FCPT_VRAD, FCPT_VAZI, FCPT_ENERGY,
    FCPT_PRESSURE, FCPT_SOUNDSPEED, FCPT_SCALE_HEIGHT, FCPT_VISCOSITY,
    FCPT_POTENTIAL, FCPT_Q_R, FCPT_Q_PHI, FCPT_DIV_V,
    FCPT_TAU_R_R, FCPT_TAU_R_PHI, FCPT_TAU_PHI_PHI, FCPT_QPLUS,
    FCPT_QMINUS, FCPT_TEMPERATURE,
    FCPT_SIGMA0, FCPT_VRAD0, FCPT_VAZI0, FCPT_ENERGY0,
    FCPT_NUM_FIELDS
} fcpt_field;

/*
 * Disk description.  Numerical and physical parameters in code units, named
 * after the reference's parameters:: / constants:: globals they stand for.
 */
typedef struct fcpt_disk_desc {
    int32_t nr;                 /* NRadial, including both ghost rings (8 .. 16384, the
                                   reference's MAX1D) */
    int32_t nphi;               /* NAzimuthal (4 .. 262144) */
    const double *radii;        /* Radii[0 .. nr+1]: nr+2 interface radii; the last one
                                   is only read through the reference's Rmed[nr] */
    double rmin, rmax;          /* configured domain edges (the reference's RMIN, RMAX
                                   globals; Radii[nr-1] may differ from RMAX in the last bit) */
    int32_t adiabatic;          /* 0 locally isothermal (AspectRatioMode 0), 1 ideal gas */
    int32_t precision;          /* FCPT_PREC_* */
    /* constants (constants.cpp: code values) */
    double G, R, sigma_sb, c_light;
    double hydro_center_mass;   /* mass the Keplerian quantities refer to */
    double mu, gamma;           /* parameters::MU, ADIABATICINDEX */
    double aspect_ratio, flaring_index; /* isothermal sound speed profile */
    double alpha;               /* parameters::viscous_alpha (CONST_ALPHA) */
    double thickness_smoothing; /* potential smoothing in units of H */
    double cfl, cfl_max_var, heating_cooling_cfl_limit;
    double sigma0, sigma_floor; /* floor = sigma_floor * sigma0 */
    double t_min, t_max;        /* temperature range (code units) */
    int32_t artificial_viscosity;  /* FCPT_ARTVISC_* */
    int32_t artvisc_dissipation;
    double artvisc_factor;
    int32_t heating_viscous; double heating_viscous_factor;
    int32_t cooling_beta_enabled; int32_t cooling_beta_ref; /* FCPT_BETA_REF_* */
    double cooling_beta, cooling_beta_ramp_up;
    double radial_viscosity_factor;
    int32_t bc_inner, bc_outer; /* FCPT_BC_*; v_phi ghost rings are Keplerian */
    int32_t damping;            /* initial-value damping of all four fields */
    double damping_inner_limit, damping_outer_limit, damping_time_factor,
        damping_time_radius_outer;
    /* last ring index damped at the inner edge / first at the outer edge, as the
       reference computes them (boundary_conditions/damping.cpp:321-327,377-383);
       -1 lets the library compute them for a logarithmic mesh */
    int32_t damp_inner_scalar, damp_inner_vector, damp_outer_scalar, damp_outer_vector;
    double omega_frame;         /* refframe::OmegaFrame (fixed frame rotation) */
    double monitor_timestep;    /* parameters::monitor_timestep for the dt clip */
    int32_t integrator;         /* FCPT_INTEGRATOR_* (parameters::hydro_integrator) */
    int32_t potential_smoothing; /* FCPT_SMOOTHING_* mask of the gas potential
                                    (CalculateNbodyPotential); 0 is the reference default */
} fcpt_disk_desc;

#define FCPT_MAX_PLANETS 8

/* Host-side N-body state consumed by one step (reference Pframeforce.cpp:27-35). */
typedef struct fcpt_planet {
    double x, y;        /* position in the hydro frame */
    double mass;        /* ramp-up mass at the current time */
    double r_cubic;     /* Klahr-Kley cubic smoothing radius; 0 disables */
} fcpt_planet;

typedef struct fcpt_step_params {
    int32_t n_planets;                 /* body 0 is the central object */
    int32_t pad_;
    fcpt_planet planet[FCPT_MAX_PLANETS];
    double indirect_x, indirect_y;     /* refframe::IndirectTerm */
} fcpt_step_params;

/* Initial condition of init_physics (init.cpp:255-343) for the profile
   setups (SigmaCondition: Profile, EnergyCondition: Profile, no self-gravity,
   no centrifugal-balance option, no random perturbation): the configuration
   keys the descriptor does not already hold. */
typedef struct fcpt_ic_desc {
    double sigma_slope;          /* SigmaSlope: Sigma = sigma0 r^-slope */
    double imposed_disk_drift;   /* ImposedDiskDrift */
    int32_t vradial_zero;        /* InitializeVradialZero */
    int32_t cutoff_outer;        /* ProfileCutoffOuter */
    int32_t cutoff_inner;        /* ProfileCutoffInner */
    int32_t pad_;
    double cutoff_outer_point, cutoff_outer_width;  /* ProfileCutoffPoint/WidthOuter */
    double cutoff_inner_point, cutoff_inner_width;  /* ProfileCutoffPoint/WidthInner */
} fcpt_ic_desc;

/* ---- context ---------------------------------------------------------- */
int fcpt_create(const fcpt_disk_desc *desc, fcpt_ctx **out);
int fcpt_destroy(fcpt_ctx *ctx);
const char *fcpt_last_error(const fcpt_ctx *ctx);
int fcpt_synchronize(fcpt_ctx *ctx);

/* ---- state transfer (host arrays in the reference layout) -------------- */
/* Sets Sigma, v_r, v_phi (and E when adiabatic), then recomputes the derived
   fields like the reference's init_euler (SourceEuler.cpp:251-284). */
int fcpt_upload_state(fcpt_ctx *ctx, const double *sigma, const double *vrad,
                      const double *vazi, const double *energy);
int fcpt_download_state(fcpt_ctx *ctx, double *sigma, double *vrad, double *vazi,
                        double *energy);
/* init_physics (init.cpp:255-343) for a profile setup, computed on the
   device: Sigma (power law, cutoffs, floor; init_gas_density, :937-1148), E
   (initial_energy with the temperature floor; init_gas_energy, :1257-1300),
   the init_euler recalculation of the derived fields and heating/cooling
   (SourceEuler.cpp:251-284, with the velocities still zero), v_phi (the
   pressure-supported smoothed Keplerian profile minus the frame rotation) and
   v_r (ImposedDiskDrift plus the numerical viscous speed, or zero;
   init_gas_velocities, :1725-1770), then copy_initial_values, the boundary
   call and copy_initial_values.  The radial profiles are formed once per ring
   on the host with the reference's expressions (O(nr)), the 2-D fields are
   filled on the device; no host arrays, no files. */
int fcpt_init_profiles(fcpt_ctx *ctx, const fcpt_ic_desc *ic);
/* Damping reference values (reference copy_initial_values, damping.cpp:301-309). */
int fcpt_upload_reference(fcpt_ctx *ctx, const double *sigma0, const double *vrad0,
                          const double *vazi0, const double *energy0);
int fcpt_upload_field(fcpt_ctx *ctx, int32_t field, const double *host);
int fcpt_download_field(fcpt_ctx *ctx, int32_t field, double *host);
int fcpt_set_time(fcpt_ctx *ctx, double time, double last_dt, int32_t n_monitor);
int fcpt_get_time(fcpt_ctx *ctx, double *time, double *last_dt, int32_t *n_monitor);
int fcpt_set_step_params(fcpt_ctx *ctx, const fcpt_step_params *params);
/* The bodies and indirect term of a LeapFrog step's second kick (the N-body
   state at midstep_time, simulation.cpp:347-361), used by fcpt_step_leapfrog,
   fcpt_run_step and fcpt_run_steps.  fcpt_set_step_params clears them: the
   second kick then sees the same bodies as the first. */
int fcpt_set_midstep_params(fcpt_ctx *ctx, const fcpt_step_params *params);

/* ---- the hydro step, one entry point per reference function ------------ */
/* cfl::condition_cfl(t_data&, double)                      cfl.h:8, cfl.cpp:185-382 */
int fcpt_condition_cfl(fcpt_ctx *ctx, double *dt_out);
/* sim::CalculateTimeStep(t_data&)               simulation.h:19, simulation.cpp:100-118 */
int fcpt_calculate_timestep(fcpt_ctx *ctx, double *dt_out);
/* CalculateNbodyPotential(t_data&, double)               Pframeforce.h, Pframeforce.cpp:21-86 */
int fcpt_nbody_potential(fcpt_ctx *ctx);
/* update_with_sourceterms(t_data&, double)             SourceEuler.h:16, SourceEuler.cpp:435-452 */
int fcpt_update_with_sourceterms(fcpt_ctx *ctx, double dt);
/* art_visc::update_with_artificial_viscosity(t_data&, double)  artificial_viscosity.cpp:11-26 */
int fcpt_artificial_viscosity(fcpt_ctx *ctx, double dt);
/* recalculate_viscosity(t_data&, double)               SourceEuler.h:12, SourceEuler.cpp:205-223 */
int fcpt_recalculate_viscosity(fcpt_ctx *ctx);
/* viscosity::compute_viscous_stress_tensor(t_data&)    viscosity.h:9, viscosity.cpp:139-350 */
int fcpt_viscous_stress_tensor(fcpt_ctx *ctx);
/* viscosity::update_velocities_with_viscosity(t_data&, double)  viscosity.cpp:355-426 */
int fcpt_update_velocities_with_viscosity(fcpt_ctx *ctx, double dt);
/* SubStep3(t_data&, double, double)                    SourceEuler.h:18, SourceEuler.cpp:859-954 */
int fcpt_substep3(fcpt_ctx *ctx, double time, double dt);
/* boundary_conditions::apply_boundary_condition(t_data&, double, double, bool)
                                          boundary_conditions.h:106, boundary_conditions.cpp:65-114 */
int fcpt_apply_boundary_condition(fcpt_ctx *ctx, double dt, int32_t final_call);
/* Transport(t_data&, PolarGrid*, PolarGrid*, PolarGrid*, PolarGrid*, double)
                                                         TransportEuler.h:5-6, TransportEuler.cpp:112-136 */
int fcpt_transport(fcpt_ctx *ctx, double dt);
/* recalculate_derived_disk_quantities(t_data&, double) SourceEuler.h:11, SourceEuler.cpp:225-249 */
int fcpt_recalculate_derived(fcpt_ctx *ctx);
/* compute_pressure(t_data&)  SourceEuler.h, SourceEuler.cpp (the pressure from
   Sigma and c_s, or from E; step_LeapFrog calls it before its second kick,
   simulation.cpp:374) */
int fcpt_compute_pressure(fcpt_ctx *ctx);
/* FARGO integer shifts of the last fcpt_transport call (TransportEuler.cpp:215-220) */
int fcpt_get_nshift(fcpt_ctx *ctx, int32_t *nshift);

/* ---- fused driver ------------------------------------------------------ */
/* step_Euler(t_data&, double) (simulation.cpp:148-267) for the gas: potential,
   sources, artificial viscosity, viscosity, SubStep3, boundaries, transport,
   final boundaries and derived fields, with the step parameters last set.
   Advances the context time by dt. */
int fcpt_step_euler(fcpt_ctx *ctx, double dt);
/* step_LeapFrog(t_data&, double) (simulation.cpp:276-459) for the gas: kick
   with dt/2 at the start time (potential of the step parameters, sources,
   artificial viscosity, viscosity, SubStep3), boundaries, transport over dt,
   kick with dt/2 at the midstep time (potential of the midstep parameters,
   compute_pressure first), final boundaries and derived fields.  Advances
   the context time by dt. */
int fcpt_step_leapfrog(fcpt_ctx *ctx, double dt);
/* One iteration of sim::run (simulation.cpp:515-556): CalculateTimeStep, the
   clip to the next monitor time, step_Euler or step_LeapFrog (the descriptor's
   integrator).  *dt_out receives the step size. */
int fcpt_run_step(fcpt_ctx *ctx, double *dt_out);

/* n_steps iterations of sim::run entirely on the device: the CFL time step,
   the monitor clip and the time stay in device memory and the planets keep the
   last step parameters.  use_graph != 0 replays one captured HIP graph per
   iteration.  No host synchronisation; errors surface at the next sync. */
int fcpt_run_steps(fcpt_ctx *ctx, int32_t n_steps, int32_t use_graph);

/* Measurement hooks (no counterpart in the reference): elapsed device time of
   n_steps fcpt_run_steps iterations, from HIP events on the context's stream;
   and the mean duration of one launch of a single kernel of the step, timed
   the same way over `reps` back-to-back launches on the current state
   (FCPT_KERNEL_* below; the listed kernels do not modify their inputs).
   FCPT_KERNEL_DISK_FORCE times the launches of fcpt_disk_on_planet_accel
   (ring means, cell reduction, block reduction; no copy to the host) for two
   bodies, a star at the origin and a planet at (1, 0), with the self-gravity
   correction.  FCPT_KERNEL_CFL is the per-ring CFL kernel (ring means of
   v_phi and the cells' limits; the fold of the ring minima is not timed).
   FCPT_KERNEL_VISC_FUSED is the fused viscous stress + velocity pass of the
   step (FCPT_PARAM_FUSED_STEP) without SubStep3: the viscous walk, or the
   cell-parallel pass where the viscous tile is 1; it writes only the
   velocities' partner buffers (and, in an adiabatic disk with viscous
   heating, the stress tensor). */
enum { FCPT_KERNEL_RADIAL_SWEEP = 0, FCPT_KERNEL_AZIMUTHAL = 1, FCPT_KERNEL_CFL = 3,
       FCPT_KERNEL_STRESS = 4, FCPT_KERNEL_POTENTIAL = 5, FCPT_KERNEL_DISK_FORCE = 6,
       FCPT_KERNEL_VISC_FUSED = 7 };
int fcpt_time_steps(fcpt_ctx *ctx, int32_t n_steps, int32_t use_graph, double *ms_total);
int fcpt_time_kernel(fcpt_ctx *ctx, int32_t kernel, int32_t reps, double *ms_per_launch);
/* Streaming-copy probe (no counterpart in the reference): copies four
   transport work arrays into their partners (8 * nr * nphi * 8 bytes moved
   per launch, returned in *bytes_per_launch) and times it like
   fcpt_time_kernel; it measures the achievable HBM bandwidth of the box and
   calibrates the profiler's byte counters.  It overwrites transport scratch
   only. */
int fcpt_time_copy(fcpt_ctx *ctx, int32_t reps, double *ms_per_launch, double *bytes_per_launch);

/* Launch-shape tuning (no counterpart in the reference).  Results do not
   depend on the shapes:
   FCPT_PARAM_RADIAL_TILE   rings walked by one thread of the radial sweep
                            (1 .. 64; 0 restores the size-based default);
   FCPT_PARAM_RADIAL_WAVES  the sweep's register budget in wavefronts per
                            SIMD (2 or 3; default 3);
   FCPT_PARAM_VISC_TILE     rings walked by one thread of the fused viscous
                            pass (1 .. 64, where 1 is the cell-parallel form;
                            0 restores the size-based default);
   FCPT_PARAM_KICK_TILE     rings walked by one thread of the kick walk, which
                            runs the source terms and the artificial
                            viscosity of the fused step as one pass (1 .. 64,
                            where 1 runs them as separate passes; 0 restores
                            the size-based default; the walk needs the TW or
                            SN artificial viscosity and a viscous tile above 1).
   FCPT_PARAM_FUSED_STEP (default 1) makes fcpt_step_euler, fcpt_step_leapfrog,
   fcpt_run_step and fcpt_run_steps run the fused passes (the walks above,
   the ghost rings with the FARGO ring means, and for Euler the final
   boundary call and derived fields inside the transport's velocity
   kernel); the evolved and
   derived fields are the same bit for bit in FP64 storage, but those steps
   leave Q_R, Q_PHI, TAU_* and DIV_V unspecified (as they were, or, for an
   adiabatic disk with viscous heating on a grid below the walks' size, the
   step's tensor); 0 runs the separate kernels of the per-function entry
   points above, which always store them.  Every parameter change drops a
   captured step graph so that the next fcpt_run_steps captures the new
   shape. */
/* FCPT_PARAM_SPLIT_ROUNDING (split storage only; default 1) chooses how a
   stored value's FP32 deviation is rounded: stochastically (up or down with
   probabilities given by the distances to the two neighbouring FP32 values, a
   deterministic hash of the value, the cell, the field and the step count
   drawing the variate: unbiased, so a slowly changing value does not lose its
   sub-unit increments every step) for the evolved fields Sigma, v_r, v_phi, E
   and the transported momenta between the two sweeps (after the second sweep
   they are FP64), to nearest for the fields derived afresh every
   step (1); stochastically for every field (2); to nearest everywhere (0). */
/* FCPT_PARAM_REBASE_INTERVAL (split storage only; default 1000) moves the
   backgrounds to the current ring means (fcpt_rebase_backgrounds) after every
   that many steps of fcpt_run_steps, fcpt_time_steps and fcpt_batch_run_steps
   (counted across calls; setting it restarts the count), so that a disk whose
   ring profiles drift (a gap opening, viscous spreading) keeps its stored
   deviations, and with them their FP32 rounding, small.  Each move reads the
   state once, copies 8 (nr + 2) bytes per evolved field to the host and
   synchronizes the stream.  0 moves them only on request.  The per-function
   step entry points (fcpt_run_step, fcpt_step_euler, ...) neither move them
   nor count; fcpt_autotune_walks' trial steps do not count either. */
enum { FCPT_PARAM_RADIAL_TILE = 0, FCPT_PARAM_RADIAL_WAVES = 1, FCPT_PARAM_FUSED_STEP = 2,
       FCPT_PARAM_VISC_TILE = 3, FCPT_PARAM_KICK_TILE = 4, FCPT_PARAM_SPLIT_ROUNDING = 5,
       FCPT_PARAM_REBASE_INTERVAL = 6 };
int fcpt_set_launch_param(fcpt_ctx *ctx, int32_t param, int32_t value);
/* The value a launch parameter takes in the next step: the size-based
   choice where it was left at 0 (tiles), else the value set. */
int fcpt_get_launch_param(fcpt_ctx *ctx, int32_t param, int32_t *value);
/* Split storage (no counterpart in the reference, whose fields are FP64):
   moves the per-ring backgrounds of Sigma, v_r, v_phi and (adiabatic) E to
   the fields' current ring means, re-expressing every deviation about its new
   background (rounded to nearest once), and re-derives the backgrounds of the
   derived fields and transport momenta from them, as an upload does.  The
   decoded values change by at most half an FP32 unit of their new deviation.
   Synchronizes the stream; FP64 storage: nothing to do. */
int fcpt_rebase_backgrounds(fcpt_ctx *ctx);
/* Autotuner (no counterpart in the reference): times every launch shape of
   the tunable kernels on the context's current state (`reps` launches each;
   the timed kernels write transport scratch only), keeps the fastest and
   returns it.  Drops a captured step graph. */
int fcpt_autotune(fcpt_ctx *ctx, int32_t reps, int32_t *radial_tile, int32_t *radial_waves);
/* Autotuner of the fused step's walks (no counterpart in the reference): the
   kick walk's and the viscous walk's tiles (2 .. 16 rings) by the time of
   `reps` whole graph-replayed iterations each, every trial from the same
   state, which the call saves first and puts back (fields and clock as they
   were); keeps the fastest and returns the tiles.  A context whose step does
   not run the walks keeps its tiles.  Drops a captured step graph. */
int fcpt_autotune_walks(fcpt_ctx *ctx, int32_t reps, int32_t *kick_tile, int32_t *visc_tile);

/* ---- batches (no counterpart in the reference) ------------------------- */
/* Many independent disks advanced together: each stage of a sim::run
   iteration is one launch over a work table holding every block of every
   disk taking part in it (disks of any size, physics and planets, on one
   device and in one storage mode, Euler integrator).  A batched disk evolves
   bit for bit like the same disk under fcpt_run_steps.  The batch uses the disks' parameters
   and launch shapes as they are at fcpt_batch_create and reads and writes
   the disks' own device state: between batch calls the disks may be read
   (fcpt_download_*, fcpt_totals) but not stepped on their own.  Every batch
   run ends with an event that each member disk's stream waits for, so such a
   read, issued right after fcpt_batch_run_steps returns, sees the stepped
   state without an explicit fcpt_batch_synchronize. */
typedef struct fcpt_batch fcpt_batch;
int fcpt_batch_create(fcpt_ctx *const *disks, int32_t n_disks, fcpt_batch **out);
int fcpt_batch_destroy(fcpt_batch *batch);
const char *fcpt_batch_last_error(const fcpt_batch *batch);
int fcpt_batch_synchronize(fcpt_batch *batch);
/* radial-sweep register budget for the whole batch (2 or 3 wavefronts per SIMD) */
int fcpt_batch_set_radial_waves(fcpt_batch *batch, int32_t waves);
/* Autotuner (no counterpart in the reference): the radial sweep's tile of the
   members that left it to the size rule (one tile for all of them, 4 .. 32
   rings) and the batch's register budget, by the time of the batched radial
   stage over `reps` launches on the current state (the timed launches write
   the transport's scratch arrays only).  Keeps and returns the fastest;
   drops a captured batch graph.  Results do not depend on the choice. */
int fcpt_batch_autotune(fcpt_batch *batch, int32_t reps, int32_t *radial_tile, int32_t *radial_waves);
/* n_steps sim::run iterations of every disk; use_graph != 0 replays one
   captured HIP graph per iteration.  No host synchronisation. */
int fcpt_batch_run_steps(fcpt_batch *batch, int32_t n_steps, int32_t use_graph);
/* as fcpt_batch_run_steps, timed with HIP events on the batch's stream */
int fcpt_batch_time_steps(fcpt_batch *batch, int32_t n_steps, int32_t use_graph, double *ms_total);
/* mean duration of one batched launch of one step kernel over all disks
   (FCPT_KERNEL_RADIAL_SWEEP, _AZIMUTHAL, _CFL, _POTENTIAL; they do not
   modify their inputs) */
int fcpt_batch_time_kernel(fcpt_batch *batch, int32_t kernel, int32_t reps, double *ms_per_launch);

/* ---- snapshots (output files of the reference) ------------------------ */
/* The run counters of a snapshot: the layout of output::misc_entry
   (output.h:16-24), the record misc.bin holds. */
typedef struct fcpt_snapshot_info {
    uint32_t snapshot;        /* sim::N_snapshot */
    uint32_t n_monitor;       /* sim::N_monitor */
    double time;              /* sim::time */
    double omega_frame;       /* refframe::OmegaFrame */
    double frame_angle;       /* refframe::FrameAngle */
    double last_dt;           /* sim::last_dt */
    uint64_t n_iter;          /* sim::N_hydro_iter */
} fcpt_snapshot_info;

/* Writes the snapshot directory `dir` (which must exist) as the reference's
   write_full_output does for the gas (output.cpp:249-305): one raw native
   FP64 file <name>.dat per field in `fields` (a mask of 1 << FCPT_*; 0
   selects Sigma, vrad, vazi, energy, and Qplus, Qminus when adiabatic), the
   whole grid including the ghost rings, vector fields with nr + 1 rows
   (t_polargrid::write2D, polargrid.cpp:135-185), and misc.bin
   (output::write_misc, output.cpp:494-525) with `info`'s counters and the
   context's time, last dt and monitor counter. */
int fcpt_write_snapshot(fcpt_ctx *ctx, const char *dir, const fcpt_snapshot_info *info, uint32_t fields);
/* restart_load for the gas (restart.cpp:19-131): misc.bin into *info and the
   context's clock, Sigma, vrad, vazi (and energy, and Qplus / Qminus when
   present, for an adiabatic disk) into the state, then the derived fields.
   Files of the wrong size are rejected. */
int fcpt_read_snapshot(fcpt_ctx *ctx, const char *dir, fcpt_snapshot_info *info);
/* The damping references from a snapshot (the reference's "reference"
   snapshot, restart.cpp:27-56): Sigma, vrad, vazi (and energy). */
int fcpt_read_snapshot_reference(fcpt_ctx *ctx, const char *dir);
/* The grid files of the output directory: used_rad.dat (Radii[0 .. nr],
   "%.18g" per line; init.cpp:227-247) and dimensions.dat (the grid header
   of parameters::write_grid_data_to_file, parameters.cpp:1075-1123, with
   `radial_spacing` e.g. "Logarithmic"). */
int fcpt_write_grid_files(fcpt_ctx *ctx, const char *outdir, const char *radial_spacing);

/* ---- the disk's gravity on the N bodies ------------------------------- */
/* One body as ComputeDiskOnPlanetAccel reads it (Force.cpp:23-122). */
typedef struct fcpt_force_body {
    double x, y;        /* position in the hydro frame */
    double r;           /* planet.get_r(): hypot(x, y) */
    double r_cubic;     /* Klahr-Kley radius l1 * CubicSmoothingFactor; 0 disables */
} fcpt_force_body;
/* compute_smoothing options (Force.cpp:145-159) and CorrectDiskSelfgravity
   (Force.cpp:64-66; the reference's default without self-gravity) */
enum { FCPT_FORCE_NO_STAR_SMOOTHING = 1, FCPT_FORCE_SMOOTHING_PLANETLOC = 2,
       FCPT_FORCE_CORRECT_SELFGRAVITY = 4 };
/* ComputeDiskOnPlanetAccel(t_data&, unsigned nb) for nb = 0 .. n_bodies-1
   (Force.h:6, Force.cpp:23-122; with FCPT_FORCE_CORRECT_SELFGRAVITY preceded
   by ComputeAverageDensity, Pframeforce.cpp:174-187, as in
   ComputeDiskOnNbodyAccel, Pframeforce.cpp:194-220): the acceleration of each
   body by the disk over the active rings, accel_xy[2k] = x, accel_xy[2k+1] = y.
   All bodies in one pass over Sigma.  Sums are formed in a fixed tree order
   (the reference's OpenMP reduction order depends on its thread count). */
int fcpt_disk_on_planet_accel(fcpt_ctx *ctx, const fcpt_force_body *bodies, int32_t n_bodies,
                              int32_t flags, double *accel_xy);

/* Compensated (Neumaier) totals over active rings with Rmed <= radius
   (quantities::gas_total_mass / gas_angular_momentum, quantities.cpp:51-78,242-279). */
int fcpt_totals(fcpt_ctx *ctx, double radius, double *mass, double *angular_momentum);

/* The ring geometry the context computed at fcpt_create (init_radialarrays,
   init.cpp:188-225): Rmed[0 .. nr-1] and Surf[0 .. nr-1] into host arrays of
   nr values (either may be NULL). */
int fcpt_get_ring_geometry(fcpt_ctx *ctx, double *rmed, double *surf);

/* No counterpart in the reference.  fcpt_totals for callers that cannot
   allocate device memory (code under HIP graph capture, tight serving loops,
   harnesses that forbid hipMalloc): the same kernels and summation order over
   caller-owned FP64 device arrays sigma and vazi (nr x nphi, the reference
   layout) with the host geometry rmed, surf (nr values each, e.g. from
   fcpt_get_ring_geometry), with no context.  The ring geometry, the per-ring
   partial sums and the results go through a static device array of each GPU
   (6 FCPT_NOALLOC_MAX_RINGS + 2 doubles, held until the process ends); the
   call allocates nothing.  Runs on `stream` (a hipStream_t, NULL for the null
   stream) and waits for it.  nr above FCPT_NOALLOC_MAX_RINGS is rejected with
   FCPT_ERR_ARG; the message is read by fcpt_noalloc_last_error().  Calls that
   overlap on one GPU share that array and race. */
#define FCPT_NOALLOC_MAX_RINGS 16384
int fcpt_totals_noalloc(int32_t nr, int32_t nphi, const double *rmed, const double *surf,
                        const double *sigma, const double *vazi, double omega_frame, double radius,
                        void *stream, double *mass, double *angular_momentum);
const char *fcpt_noalloc_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* FCPT_H */
