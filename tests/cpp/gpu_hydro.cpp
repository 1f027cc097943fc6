// gpu_hydro.cpp -- the reference-side adapter that drops libfcpt_hip.so into
// FargoCPT: a new file next to src/simulation.cpp in the reference tree,
// compiled against the reference's own headers and include/fcpt.h.  It keeps
// the reference's call sites (simulation.cpp:100-267, Pframeforce.cpp:194-220,
// output.cpp:249-305) and turns them into calls of the C ABI; the gas state
// stays in HBM between calls and is copied to t_data's host arrays only where
// the reference reads them outside the hydro step (output, monitor).
//
// tests/test_adapter.py compiles this file against /root/reference/src and
// links it with the reference's objects and libfcpt_hip.so, so a renamed
// parameters:: / boundary_conditions:: symbol or a changed fcpt_* signature
// fails the build.
#include <cstdio>
#include <string>

#include "fcpt.h"

#include "LowTasks.h"
#include "Theo.h"
#include "boundary_conditions/boundary_conditions.h"
#include "constants.h"
#include "data.h"
#include "find_cell_id.h"
#include "frame_of_reference.h"
#include "global.h"
#include "output.h"
#include "parameters.h"
#include "simulation.h"
#include "types.h"

namespace gpu_hydro
{
fcpt_ctx *ctx = nullptr;

static void check(int rc)
{
    if (rc) die("fcpt: %s\n", fcpt_last_error(ctx));
}

static int32_t bc_of(const std::string &vrad_name)
{
    // the composite boundaries the library implements (v_r decides them,
    // boundary_conditions/config.cpp:193-258; Sigma and E zero-gradient,
    // v_phi Keplerian)
    if (vrad_name == "reflecting") return FCPT_BC_REFLECTING;
    if (vrad_name == "zerogradient") return FCPT_BC_ZEROGRADIENT;
    if (vrad_name == "outflow") return FCPT_BC_OUTFLOW;
    die("gpu_hydro: boundary '%s' is not on the GPU path\n", vrad_name.c_str());
    return -1;
}

// after init_physics() (main.cpp:116): the context and the initial state
void init(t_data &data, int32_t precision)
{
    fcpt_disk_desc d{};
    d.nr = (int32_t)NRadial;
    d.nphi = (int32_t)NAzimuthal;
    d.radii = &Radii[0];  // Radii[0 .. NRadial + 1]
    d.rmin = RMIN;
    d.rmax = RMAX;
    d.adiabatic = parameters::Adiabatic ? 1 : 0;
    d.precision = precision;  // FCPT_PREC_FP64 or FCPT_PREC_SPLIT32
    d.G = constants::G.get_code_value();
    d.R = constants::R.get_code_value();
    d.sigma_sb = constants::sigma.get_code_value();
    d.c_light = constants::c.get_code_value();
    d.hydro_center_mass = hydro_center_mass;
    d.mu = parameters::MU;
    d.gamma = parameters::ADIABATICINDEX;
    d.aspect_ratio = parameters::aspectratio_ref;
    d.flaring_index = parameters::flaring_index;
    d.alpha = parameters::viscous_alpha;
    d.thickness_smoothing = parameters::thickness_smoothing;
    d.cfl = parameters::CFL;
    d.cfl_max_var = parameters::CFL_max_var;
    d.heating_cooling_cfl_limit = parameters::HEATING_COOLING_CFL_LIMIT;
    d.sigma0 = parameters::sigma0;
    d.sigma_floor = parameters::sigma_floor;
    d.t_min = parameters::minimum_temperature;
    d.t_max = parameters::maximum_temperature;
    d.artificial_viscosity = (int32_t)parameters::artificial_viscosity;  // none / TW / SN
    d.artvisc_dissipation = parameters::artificial_viscosity_dissipation ? 1 : 0;
    d.artvisc_factor = parameters::artificial_viscosity_factor;
    d.heating_viscous = parameters::heating_viscous_enabled ? 1 : 0;
    d.heating_viscous_factor = parameters::heating_viscous_factor;
    d.cooling_beta_enabled = parameters::cooling_beta_enabled ? 1 : 0;
    d.cooling_beta_ref = parameters::cooling_beta_floor       ? FCPT_BETA_REF_FLOOR
                         : parameters::cooling_beta_reference ? FCPT_BETA_REF_REFERENCE
                                                              : FCPT_BETA_REF_ZERO;
    d.cooling_beta = parameters::cooling_beta;
    d.cooling_beta_ramp_up = parameters::cooling_beta_ramp_up;
    d.radial_viscosity_factor = parameters::radial_viscosity_factor;
    d.bc_inner = bc_of(boundary_conditions::vrad_inner_name);
    d.bc_outer = bc_of(boundary_conditions::vrad_outer_name);
    d.damping = boundary_conditions::damping_enabled ? 1 : 0;
    d.damping_inner_limit = boundary_conditions::damping_inner_limit;
    d.damping_outer_limit = boundary_conditions::damping_outer_limit;
    d.damping_time_factor = boundary_conditions::damping_time_factor;
    d.damping_time_radius_outer = boundary_conditions::damping_time_radius_outer;
    // the damped rings as damping.cpp:321-327, 377-383 find them
    d.damp_inner_scalar = (int32_t)clamp_r_id_to_rmed_grid(
        get_rmed_id(RMIN * boundary_conditions::damping_inner_limit), false);
    d.damp_inner_vector = (int32_t)clamp_r_id_to_radii_grid(
        get_rinf_id(RMIN * boundary_conditions::damping_inner_limit), true);
    d.damp_outer_scalar = (int32_t)clamp_r_id_to_rmed_grid(
        get_rmed_id(RMAX * boundary_conditions::damping_outer_limit) + 1, false);
    d.damp_outer_vector = (int32_t)clamp_r_id_to_radii_grid(
        get_rinf_id(RMAX * boundary_conditions::damping_outer_limit) + 1, true);
    d.omega_frame = refframe::OmegaFrame;
    d.monitor_timestep = parameters::monitor_timestep;
    d.integrator = parameters::hydro_integrator == LEAPFROG_INTEGRATOR ? FCPT_INTEGRATOR_LEAPFROG
                                                                       : FCPT_INTEGRATOR_EULER;
    d.potential_smoothing =
        (parameters::compatibility_no_star_smoothing ? FCPT_SMOOTHING_NO_STAR : 0) |
        (parameters::compatibility_smoothing_planetloc ? FCPT_SMOOTHING_PLANETLOC : 0);
    check(fcpt_create(&d, &ctx));
    int32_t tile = 0, waves = 0;
    check(fcpt_autotune(ctx, 10, &tile, &waves));  // launch shapes for this grid (optional)
    check(fcpt_upload_reference(ctx, data[t_data::SIGMA0].Field, data[t_data::V_RADIAL0].Field,
                                data[t_data::V_AZIMUTHAL0].Field, data[t_data::ENERGY0].Field));
    check(fcpt_upload_state(ctx, data[t_data::SIGMA].Field, data[t_data::V_RADIAL].Field,
                            data[t_data::V_AZIMUTHAL].Field,
                            parameters::Adiabatic ? data[t_data::ENERGY].Field : nullptr));
}

// a profile setup initialised on the device instead of init_physics + upload
void init_on_device()
{
    fcpt_ic_desc ic{};
    ic.sigma_slope = parameters::sigma_slope;
    ic.imposed_disk_drift = parameters::IMPOSEDDISKDRIFT;
    ic.vradial_zero = parameters::initialize_vradial_zero ? 1 : 0;
    ic.cutoff_outer = parameters::profile_cutoff_outer ? 1 : 0;
    ic.cutoff_outer_point = parameters::profile_cutoff_point_outer;
    ic.cutoff_outer_width = parameters::profile_cutoff_width_outer;
    ic.cutoff_inner = parameters::profile_cutoff_inner ? 1 : 0;
    ic.cutoff_inner_point = parameters::profile_cutoff_point_inner;
    ic.cutoff_inner_width = parameters::profile_cutoff_width_inner;
    check(fcpt_init_profiles(ctx, &ic));
}

// the N-body module's per-step outputs (Pframeforce.cpp:27-35) to the device
static fcpt_step_params bodies_at(t_data &data, double t)
{
    t_planetary_system &ps = data.get_planetary_system();
    fcpt_step_params p{};
    p.n_planets = (int32_t)ps.get_number_of_planets();
    if (p.n_planets > FCPT_MAX_PLANETS) die("gpu_hydro: at most %d bodies\n", FCPT_MAX_PLANETS);
    for (int k = 0; k < p.n_planets; ++k) {
        t_planet &pl = ps.get_planet(k);
        p.planet[k].x = pl.get_x();
        p.planet[k].y = pl.get_y();
        p.planet[k].mass = pl.get_rampup_mass(t);
        p.planet[k].r_cubic = pl.get_dimensionless_roche_radius() * pl.get_distance_to_primary() *
                              pl.get_cubic_smoothing_factor();
    }
    p.indirect_x = refframe::IndirectTerm.x;
    p.indirect_y = refframe::IndirectTerm.y;
    return p;
}

void publish_bodies(t_data &data, double t)
{
    const fcpt_step_params p = bodies_at(data, t);
    check(fcpt_set_step_params(ctx, &p));
}

// before snapshots and monitor output: the state to t_data's host arrays
void download(t_data &data)
{
    check(fcpt_download_state(ctx, data[t_data::SIGMA].Field, data[t_data::V_RADIAL].Field,
                              data[t_data::V_AZIMUTHAL].Field,
                              parameters::Adiabatic ? data[t_data::ENERGY].Field : nullptr));
}

// sim::CalculateTimeStep (simulation.cpp:100-118)
double calculate_time_step(t_data &data)
{
    (void)data;
    double rv = sim::last_dt;
    if (parameters::calculate_disk) {
        check(fcpt_set_time(ctx, sim::time, sim::last_dt, (int32_t)sim::N_monitor));
        check(fcpt_calculate_timestep(ctx, &rv));  // condition_cfl + the CFLmaxVar limit
        sim::last_dt = rv;
    }
    return rv;
}

// the gas part of step_Euler (simulation.cpp:148-267), one entry per
// reference call; the N-body calls around it stay on the host
void step_euler_gas_before_transport(t_data &data, double dt)
{
    publish_bodies(data, sim::time);                        // g_xpl / g_ypl / g_mpl set-up
    check(fcpt_nbody_potential(ctx));                       // CalculateNbodyPotential
    check(fcpt_update_with_sourceterms(ctx, dt));           // update_with_sourceterms
    check(fcpt_artificial_viscosity(ctx, dt));              // art_visc::update_with_artificial_viscosity
    check(fcpt_recalculate_viscosity(ctx));                 // recalculate_viscosity
    check(fcpt_viscous_stress_tensor(ctx));                 // viscosity::compute_viscous_stress_tensor
    check(fcpt_update_velocities_with_viscosity(ctx, dt));  // viscosity::update_velocities_with_viscosity
    if (parameters::Adiabatic) check(fcpt_substep3(ctx, sim::time, dt));  // SubStep3
    check(fcpt_apply_boundary_condition(ctx, 0.0, 0));      // apply_boundary_condition(.., false)
    check(fcpt_transport(ctx, dt));                         // Transport
}
void step_euler_gas_after_transport(double dt)
{
    // after the N-body integration and time += dt
    check(fcpt_apply_boundary_condition(ctx, dt, 1));       // apply_boundary_condition(.., true)
    check(fcpt_recalculate_derived(ctx));                   // recalculate_derived_disk_quantities
}

// the whole sim::run iteration on the device, for bodies on fixed orbits
double run_iteration()
{
    double dt = 0.0;
    check(fcpt_run_step(ctx, &dt));
    return dt;
}

// ComputeDiskOnNbodyAccel (Pframeforce.cpp:194-220): all bodies in one pass
void compute_disk_on_nbody_accel(t_data &data)
{
    t_planetary_system &ps = data.get_planetary_system();
    const int n = (int)ps.get_number_of_planets();
    if (n > FCPT_MAX_PLANETS) die("gpu_hydro: at most %d bodies\n", FCPT_MAX_PLANETS);
    fcpt_force_body b[FCPT_MAX_PLANETS] = {};
    for (int k = 0; k < n; ++k) {
        t_planet &pl = ps.get_planet(k);
        const double l1 = pl.get_dimensionless_roche_radius() * pl.get_distance_to_primary();
        const double c = pl.get_cubic_smoothing_factor();
        b[k] = fcpt_force_body{pl.get_x(), pl.get_y(), pl.get_r(), c > 0.0 ? l1 * c : 0.0};
    }
    const int flags =
        (parameters::correct_disk_selfgravity ? FCPT_FORCE_CORRECT_SELFGRAVITY : 0) |
        (parameters::compatibility_no_star_smoothing ? FCPT_FORCE_NO_STAR_SMOOTHING : 0) |
        (parameters::compatibility_smoothing_planetloc ? FCPT_FORCE_SMOOTHING_PLANETLOC : 0);
    double acc[2 * FCPT_MAX_PLANETS];
    check(fcpt_disk_on_planet_accel(ctx, b, n, flags, acc));
    for (int k = 0; k < n; ++k) {
        t_planet &pl = ps.get_planet(k);
        Pair a;
        a.x = acc[2 * k];
        a.y = acc[2 * k + 1];
        pl.set_disk_on_planet_acceleration(a);
        pl.set_torque((pl.get_x() * a.y - pl.get_y() * a.x) * pl.get_mass());
    }
}

// write_full_output for the gas grids (output.cpp:249-305)
void write_snapshot(uint32_t fields)
{
    fcpt_snapshot_info info{};
    info.snapshot = sim::N_snapshot;
    info.omega_frame = refframe::OmegaFrame;
    info.frame_angle = refframe::FrameAngle;
    info.n_iter = sim::N_hydro_iter;
    check(fcpt_write_snapshot(ctx, output::snapshot_dir.c_str(), &info, fields));
}

// restart_load for the gas (restart.cpp:19-131)
void read_snapshot(const std::string &dir, const std::string &reference_dir)
{
    fcpt_snapshot_info info{};
    check(fcpt_read_snapshot(ctx, dir.c_str(), &info));
    check(fcpt_read_snapshot_reference(ctx, reference_dir.c_str()));
    sim::time = info.time;
    sim::last_dt = info.last_dt;
    sim::N_monitor = info.n_monitor;
    sim::N_snapshot = info.snapshot;
    sim::N_hydro_iter = info.n_iter;
}

void finish()
{
    fcpt_destroy(ctx);
    ctx = nullptr;
}
}  // namespace gpu_hydro
