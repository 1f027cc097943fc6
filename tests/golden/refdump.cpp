// Fixture driver for the reference FargoCPT hydro step.
//
// Links the reference's own object files (built from /root/reference/src by
// oracle/ref.mk into oracle/_ref/) and drives them exactly like the reference's
// main() (src/main.cpp:48-164) and sim::run / step_Euler (src/simulation.cpp:
// 148-267, 462-558), so that every array written here is produced by the
// reference's functions.  Two deviations, both for determinism and both
// applied before the first step:
//   * the initial gas state is replaced by arrays read from --ic (written by
//     tests/golden/make_golden.py from the package's seeded initial condition),
//     followed by the reference's own init_euler recalculation of the derived
//     fields (src/SourceEuler.cpp:251-284) and copy of the damping reference
//     values (src/boundary_conditions/damping.cpp:300-309);
//   * logging starts after MPI is initialised (the reference's
//     register_signal_handlers() logs before CPU_Number is set, which
//     crashes a build without -ffast-math, src/logging.cpp:132).
//
// Output directory layout (all binary arrays are native little-endian):
//   meta.txt                key = value lines (constants, sizes, damping ids)
//   geom_<name>.bin         1-D radial arrays (float64)
//   steps.bin               per step: 16 float64 values (see STEP_RECORD)
//   planets.bin             per step: n_planets x 4 float64 (x, y, m, r_cubic), twice
//                           per step for LeapFrog (first kick, then second kick)
//   nshift.bin              per step: Nr int32 FARGO shifts
//   diskacc.bin             per step, after it: n_planets x 6 float64 (x, y, r, Klahr
//                           radius, ax, ay) of ComputeDiskOnPlanetAccel, one thread
//   totals.bin              per step, after it: 5 float64 (gas_total_mass and
//                           gas_angular_momentum over the whole disk, the same two
//                           up to the radius Rmed[Nr/2], and that radius), one thread
//   state_<step>_<f>.bin    checkpoints (f = sigma, vrad, vazi, energy, qplus, qminus);
//                           state_ic_<f> and ic_<derived>.bin: the reference's own
//                           initial condition when no --ic directory is given
//   stage_<step>_<k>_<name>_<f>.bin   per-stage fields for the first steps
//   snapshot/               with REFDUMP_SNAPSHOT set: the reference's snapshot
//                           files (<grid>.dat, misc.bin) of the final state
// With REFDUMP_TIME set the step loop writes nothing per step and meta.txt gets
// loop_seconds, the wall time of the nsteps sim::run iterations (the CPU
// baseline's fairness check times the reference this way).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <set>

#include <omp.h>

#include "Force.h"
#include "Interpret.h"
#include "LowTasks.h"
#include "Pframeforce.h"
#include "SourceEuler.h"
#include "Theo.h"
#include "TransportEuler.h"
#include "boundary_conditions/boundary_conditions.h"
#include "cfl.h"
#include "commbound.h"
#include "config.h"
#include "constants.h"
#include "data.h"
#include "find_cell_id.h"
#include "frame_of_reference.h"
#include "global.h"
#include "init.h"
#include "options.h"
#include "output.h"
#include "parallel.h"
#include "parameters.h"
#include "random/random_wrapper.h"
#include "simulation.h"
#include "split.h"
#include "accretion.h"
#include "quantities.h"
#include "units.h"
#include "viscosity/artificial_viscosity.h"
#include "viscosity/viscosity.h"

namespace boundary_conditions {
extern double damping_inner_limit, damping_outer_limit, damping_time_factor,
    damping_time_radius_outer;
}

static std::string g_out;
static const int *g_nshift = nullptr;

static void write_raw(const std::string &name, const void *p, size_t bytes)
{
    const std::string path = g_out + "/" + name;
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) {
	fprintf(stderr, "cannot write %s\n", path.c_str());
	exit(2);
    }
    fwrite(p, 1, bytes, f);
    fclose(f);
}

static void append_raw(const std::string &name, const void *p, size_t bytes)
{
    const std::string path = g_out + "/" + name;
    FILE *f = fopen(path.c_str(), "ab");
    fwrite(p, 1, bytes, f);
    fclose(f);
}

static void read_raw(const std::string &path, double *p, size_t n)
{
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) {
	fprintf(stderr, "cannot read %s\n", path.c_str());
	exit(2);
    }
    if (fread(p, sizeof(double), n, f) != n) {
	fprintf(stderr, "short read %s\n", path.c_str());
	exit(2);
    }
    fclose(f);
}

static size_t grid_count(t_polargrid &g)
{
    return (size_t)g.get_size_radial() * (size_t)g.get_size_azimuthal();
}

static void dump_grid(const std::string &name, t_polargrid &g)
{
    write_raw(name + ".bin", g.Field, grid_count(g) * sizeof(double));
}

struct Named {
    const char *name;
    t_data::t_polargrid_type id;
};

static const Named kStageFields[] = {
    {"sigma", t_data::SIGMA},        {"vrad", t_data::V_RADIAL},
    {"vazi", t_data::V_AZIMUTHAL},   {"energy", t_data::ENERGY},
    {"pressure", t_data::PRESSURE},  {"soundspeed", t_data::SOUNDSPEED},
    {"scaleheight", t_data::SCALE_HEIGHT}, {"viscosity", t_data::VISCOSITY},
    {"potential", t_data::POTENTIAL}, {"qr", t_data::Q_R},
    {"qphi", t_data::Q_PHI},         {"divv", t_data::DIV_V},
    {"taurr", t_data::TAU_R_R},      {"taurp", t_data::TAU_R_PHI},
    {"taupp", t_data::TAU_PHI_PHI},  {"qplus", t_data::QPLUS},
    {"qminus", t_data::QMINUS},      {"temperature", t_data::TEMPERATURE},
};

static void dump_stage(t_data &data, unsigned step, int k, const char *stage)
{
    char prefix[128];
    snprintf(prefix, sizeof(prefix), "stage_%u_%02d_%s_", step, k, stage);
    for (const Named &n : kStageFields) {
	dump_grid(std::string(prefix) + n.name, data[n.id]);
    }
}

static void dump_state(t_data &data, const std::string &tag)
{
    dump_grid("state_" + tag + "_sigma", data[t_data::SIGMA]);
    dump_grid("state_" + tag + "_vrad", data[t_data::V_RADIAL]);
    dump_grid("state_" + tag + "_vazi", data[t_data::V_AZIMUTHAL]);
    dump_grid("state_" + tag + "_energy", data[t_data::ENERGY]);
    dump_grid("state_" + tag + "_qplus", data[t_data::QPLUS]);
    dump_grid("state_" + tag + "_qminus", data[t_data::QMINUS]);
}

// Same order of calls as step_Euler (src/simulation.cpp:148-267) for the
// physics enabled by the fixture configs (no accretion onto planets, no disk
// feedback, no particles, no radiative diffusion, no self-gravity).
static void euler_step(t_data &data, double dt, unsigned step, bool stages,
		       std::vector<double> &planet_rec, double *ind_term)
{
    int k = 0;
    if (parameters::calculate_disk) {
	accretion::AccreteOntoPlanets(data, dt);
    }
    if (parameters::disk_feedback) {
	ComputeDiskOnNbodyAccel(data);
	UpdatePlanetVelocitiesWithDiskForce(data, dt);
    }
    refframe::ComputeIndirectTermDisk(data, dt);
    refframe::ComputeIndirectTermNbody(data, sim::time, dt);
    refframe::ComputeIndirectTermFully();
    data.get_planetary_system().apply_indirect_term_on_Nbody(
	refframe::IndirectTerm, dt);

    if (stages) dump_stage(data, step, k++, "start");
    CalculateNbodyPotential(data, sim::time);
    // record the host-side inputs the potential consumed
    const unsigned np = data.get_planetary_system().get_number_of_planets();
    for (unsigned p = 0; p < np; ++p) {
	planet_rec.push_back(g_xpl[p]);
	planet_rec.push_back(g_ypl[p]);
	planet_rec.push_back(g_mpl[p]);
	planet_rec.push_back(g_cubic_smoothing_radius[p]);
    }
    ind_term[0] = refframe::IndirectTerm.x;
    ind_term[1] = refframe::IndirectTerm.y;
    if (stages) dump_stage(data, step, k++, "potential");

    refframe::handle_corotation(data, dt);

    update_with_sourceterms(data, dt);
    if (stages) dump_stage(data, step, k++, "sources");
    art_visc::update_with_artificial_viscosity(data, dt);
    if (stages) dump_stage(data, step, k++, "artvisc");
    recalculate_viscosity(data, sim::time);
    if (stages) dump_stage(data, step, k++, "viscosity");
    viscosity::compute_viscous_stress_tensor(data);
    if (stages) dump_stage(data, step, k++, "stress");
    viscosity::update_velocities_with_viscosity(data, dt);
    if (stages) dump_stage(data, step, k++, "viscupdate");
    if (parameters::Adiabatic) {
	SubStep3(data, sim::time, dt);
	if (stages) dump_stage(data, step, k++, "substep3");
    }
    boundary_conditions::apply_boundary_condition(data, sim::time, 0.0, false);
    if (stages) dump_stage(data, step, k++, "bc");
    Transport(data, &data[t_data::SIGMA], &data[t_data::V_RADIAL],
	      &data[t_data::V_AZIMUTHAL], &data[t_data::ENERGY], dt);
    if (g_nshift) {
	append_raw("nshift.bin", g_nshift, sizeof(int) * NRadial);
    }
    if (stages) dump_stage(data, step, k++, "transport");

    data.get_planetary_system().integrate(sim::time, dt);
    data.get_planetary_system().copy_data_from_rebound();
    data.get_planetary_system()
	.move_to_hydro_center_and_update_orbital_parameters();

    sim::time += dt;
    sim::N_hydro_iter = sim::N_hydro_iter + 1;

    CommunicateBoundaries(&data[t_data::SIGMA], &data[t_data::V_RADIAL],
			  &data[t_data::V_AZIMUTHAL], &data[t_data::ENERGY]);
    boundary_conditions::apply_boundary_condition(data, sim::time, dt, true);
    if (stages) dump_stage(data, step, k++, "bcfinal");
    recalculate_derived_disk_quantities(data, sim::time);
    if (stages) dump_stage(data, step, k++, "derived");
}

// Same order of calls as step_LeapFrog (src/simulation.cpp:276-459) for the
// same physics: gas kick with dt/2 at the start positions, drift (transport
// over dt), second kick with dt/2 at the midstep positions, the N-body drift
// and kick halves around them.  The potential inputs of both kicks are
// recorded (planet_rec holds the first kick's bodies, then the second's;
// ind_term[0..1] and ind_term[4..5] the two indirect terms).
static void leapfrog_step(t_data &data, const double step_dt, unsigned step, bool stages,
			  std::vector<double> &planet_rec, double *ind_term)
{
    int k = 0;
    const double frog_dt = step_dt / 2;
    const double start_time = sim::time;
    const double midstep_time = sim::time + frog_dt;
    const double end_time = sim::time + step_dt;
    const unsigned np = data.get_planetary_system().get_number_of_planets();
    auto record = [&](double *ind) {
	for (unsigned p = 0; p < np; ++p) {
	    planet_rec.push_back(g_xpl[p]);
	    planet_rec.push_back(g_ypl[p]);
	    planet_rec.push_back(g_mpl[p]);
	    planet_rec.push_back(g_cubic_smoothing_radius[p]);
	}
	ind[0] = refframe::IndirectTerm.x;
	ind[1] = refframe::IndirectTerm.y;
    };

    refframe::ComputeIndirectTermNbody(data, start_time, frog_dt);
    refframe::init_corotation(data);
    data.get_planetary_system().integrate(start_time, frog_dt);
    data.get_planetary_system().copy_data_from_rebound();
    data.get_planetary_system().move_to_hydro_center_and_update_orbital_parameters();
    if (parameters::disk_feedback) {
	ComputeDiskOnNbodyAccel(data);
    }
    refframe::ComputeIndirectTermDisk(data, frog_dt);
    refframe::ComputeIndirectTermFully();
    accretion::AccreteOntoPlanets(data, frog_dt);
    if (parameters::disk_feedback) {
	UpdatePlanetVelocitiesWithDiskForce(data, frog_dt);
    }
    data.get_planetary_system().apply_indirect_term_on_Nbody(refframe::IndirectTerm, frog_dt);
    refframe::handle_corotation(data, frog_dt);

    if (stages) dump_stage(data, step, k++, "start");
    CalculateNbodyPotential(data, start_time);
    record(&ind_term[0]);
    if (stages) dump_stage(data, step, k++, "potential");
    update_with_sourceterms(data, frog_dt);
    if (stages) dump_stage(data, step, k++, "sources");
    art_visc::update_with_artificial_viscosity(data, frog_dt);
    if (stages) dump_stage(data, step, k++, "artvisc");
    recalculate_viscosity(data, start_time);
    viscosity::compute_viscous_stress_tensor(data);
    viscosity::update_velocities_with_viscosity(data, frog_dt);
    if (stages) dump_stage(data, step, k++, "viscupdate");
    if (parameters::Adiabatic) {
	SubStep3(data, start_time, frog_dt);
	if (stages) dump_stage(data, step, k++, "substep3");
    }
    boundary_conditions::apply_boundary_condition(data, start_time, 0.0, false);
    Transport(data, &data[t_data::SIGMA], &data[t_data::V_RADIAL],
	      &data[t_data::V_AZIMUTHAL], &data[t_data::ENERGY], step_dt);
    if (g_nshift) {
	append_raw("nshift.bin", g_nshift, sizeof(int) * NRadial);
    }
    if (stages) dump_stage(data, step, k++, "transport");

    if (parameters::disk_feedback) {
	ComputeDiskOnNbodyAccel(data);
    }
    refframe::ComputeIndirectTermDisk(data, frog_dt);
    refframe::ComputeIndirectTermNbody(data, midstep_time, frog_dt);
    refframe::ComputeIndirectTermFully();
    CalculateNbodyPotential(data, midstep_time);
    record(&ind_term[4]);
    if (stages) dump_stage(data, step, k++, "potential2");
    compute_pressure(data);
    update_with_sourceterms(data, frog_dt);
    if (stages) dump_stage(data, step, k++, "sources2");
    art_visc::update_with_artificial_viscosity(data, frog_dt);
    if (stages) dump_stage(data, step, k++, "artvisc2");
    recalculate_viscosity(data, midstep_time);
    viscosity::compute_viscous_stress_tensor(data);
    viscosity::update_velocities_with_viscosity(data, frog_dt);
    if (stages) dump_stage(data, step, k++, "viscupdate2");
    if (parameters::Adiabatic) {
	SubStep3(data, midstep_time, frog_dt);
	if (stages) dump_stage(data, step, k++, "substep3_2");
    }

    accretion::AccreteOntoPlanets(data, frog_dt);
    if (parameters::disk_feedback) {
	UpdatePlanetVelocitiesWithDiskForce(data, frog_dt);
    }
    data.get_planetary_system().apply_indirect_term_on_Nbody(refframe::IndirectTerm, frog_dt);
    refframe::init_corotation(data);
    data.get_planetary_system().integrate(midstep_time, frog_dt);
    data.get_planetary_system().copy_data_from_rebound();
    data.get_planetary_system().move_to_hydro_center_and_update_orbital_parameters();
    refframe::handle_corotation(data, frog_dt);

    sim::time = end_time;
    sim::N_hydro_iter = sim::N_hydro_iter + 1;
    CommunicateBoundaries(&data[t_data::SIGMA], &data[t_data::V_RADIAL],
			  &data[t_data::V_AZIMUTHAL], &data[t_data::ENERGY]);
    boundary_conditions::apply_boundary_condition(data, end_time, step_dt, true);
    if (stages) dump_stage(data, step, k++, "bcfinal");
    recalculate_derived_disk_quantities(data, end_time);
    if (stages) dump_stage(data, step, k++, "derived");
}

// The disk's gravity on every body after a step, as ComputeDiskOnNbodyAccel
// obtains it (Pframeforce.cpp:194-220), with one OpenMP thread so that the
// reduction runs in the loop order; the planets' own state is not touched.
static void dump_disk_accel(t_data &data)
{
    const int nt = omp_get_max_threads();
    omp_set_num_threads(1);
    if (parameters::correct_disk_selfgravity) {
	ComputeAverageDensity(data);
    }
    std::vector<double> rec;
    t_planetary_system &psys = data.get_planetary_system();
    for (unsigned k = 0; k < psys.get_number_of_planets(); ++k) {
	t_planet &pl = psys.get_planet(k);
	const double c = pl.get_cubic_smoothing_factor();
	const double rsm = c > 0.0 ? pl.get_dimensionless_roche_radius() *
					 pl.get_distance_to_primary() * c
				   : 0.0;
	const Pair a = ComputeDiskOnPlanetAccel(data, k);
	rec.push_back(pl.get_x());
	rec.push_back(pl.get_y());
	rec.push_back(pl.get_r());
	rec.push_back(rsm);
	rec.push_back(a.x);
	rec.push_back(a.y);
    }
    omp_set_num_threads(nt);
    append_raw("diskacc.bin", rec.data(), sizeof(double) * rec.size());
}

// The gas totals after a step (quantities::gas_total_mass,
// quantities.cpp:51-78, and quantities::gas_angular_momentum, :242-279), with
// one OpenMP thread so that each sum runs in the loop order, over the whole
// disk and up to the radius of the middle ring (Rmed <= radius includes that
// ring itself).
static void dump_totals(t_data &data)
{
    const int nt = omp_get_max_threads();
    omp_set_num_threads(1);
    const double r_cut = Rmed[data[t_data::SIGMA].get_size_radial() / 2];
    double rec[5];
    rec[0] = quantities::gas_total_mass(data, 1e300);
    rec[1] = quantities::gas_angular_momentum(data, 1e300);
    rec[2] = quantities::gas_total_mass(data, r_cut);
    rec[3] = quantities::gas_angular_momentum(data, r_cut);
    rec[4] = r_cut;
    omp_set_num_threads(nt);
    append_raw("totals.bin", rec, sizeof(rec));
}

int main(int argc, char **argv)
{
    // refdump <config.yml> <outdir> <icdir|-> <nsteps> <stage_steps> <checkpoints,..> [nshift_addr|0] [simrun]
    if (argc < 8) {
	fprintf(stderr, "usage: %s config out icdir nsteps stage_steps checkpoints nshift_addr [simrun]\n", argv[0]);
	return 2;
    }
    const std::string cfgfile = argv[1];
    g_out = argv[2];
    const std::string icdir = argv[3];
    const unsigned nsteps = (unsigned)atoi(argv[4]);
    const unsigned stage_steps = (unsigned)atoi(argv[5]);
    std::set<unsigned> checkpoints;
    {
	std::string s = argv[6];
	size_t pos = 0;
	while (pos < s.size()) {
	    size_t e = s.find(',', pos);
	    if (e == std::string::npos) e = s.size();
	    if (e > pos) checkpoints.insert((unsigned)atoi(s.substr(pos, e - pos).c_str()));
	    pos = e + 1;
	}
    }
    const unsigned long addr = strtoul(argv[7], nullptr, 0);
    if (addr) g_nshift = reinterpret_cast<const int *>(addr);
    const bool simrun = argc > 8 && std::string(argv[8]) == "simrun";

    // options::parse expects the reference command line: "start <config>"
    char arg0[] = "fargocpt", arg1[] = "start";
    std::vector<char> arg2(cfgfile.begin(), cfgfile.end());
    arg2.push_back('\0');
    char *fargv[] = {arg0, arg1, arg2.data(), nullptr};
    int fargc = 3;
    init_parallel(fargc, fargv);
    options::parse(fargc, fargv);

    t_data data;
    resize_radialarrays(MAX1D);
    ReadVariables(options::parameter_file, data, fargc, fargv);
    parameters::write_grid_data_to_file();
    units::write_code_units_file();
    constants::write_code_constants_file();
    SplitDomain();
    data.set_size(GlobalNRadial, NAzimuthal, NRadial, NAzimuthal);
    fargo_random::init();
    init_radialarrays();
    data.get_planetary_system().init_system();
    data.get_massflow_tracker().init(data.get_planetary_system());
    init_binary_quadropole_moment(data.get_planetary_system());
    // also decides whether damping reference values are stored
    // (boundary_conditions::describe_damping, src/parameters.cpp:952)
    parameters::summarize_parameters();
    config::cfg.exit_on_unknown_key();
    boundary_conditions::init(data);
    init_physics(data);
    if (icdir == "-") {
	// the reference's own initial condition (init_physics, src/init.cpp:255-343)
	dump_state(data, "ic");
	dump_grid("ic_pressure", data[t_data::PRESSURE]);
	dump_grid("ic_soundspeed", data[t_data::SOUNDSPEED]);
	dump_grid("ic_temperature", data[t_data::TEMPERATURE]);
	dump_grid("ic_viscosity", data[t_data::VISCOSITY]);
    }

    const unsigned Nr = NRadial, Ns = NAzimuthal;
    if (icdir != "-") {
	read_raw(icdir + "/sigma.bin", data[t_data::SIGMA].Field, (size_t)Nr * Ns);
	read_raw(icdir + "/vrad.bin", data[t_data::V_RADIAL].Field, (size_t)(Nr + 1) * Ns);
	read_raw(icdir + "/vazi.bin", data[t_data::V_AZIMUTHAL].Field, (size_t)Nr * Ns);
	read_raw(icdir + "/energy.bin", data[t_data::ENERGY].Field, (size_t)Nr * Ns);
	// the reference's init_euler recalculation (src/SourceEuler.cpp:264-283)
	if (parameters::Locally_Isothermal) {
	    compute_sound_speed(data, sim::time);
	    compute_pressure(data);
	    compute_temperature(data);
	    compute_scale_height(data, sim::time);
	}
	if (parameters::Adiabatic) {
	    compute_temperature(data);
	    compute_sound_speed(data, sim::time);
	    compute_scale_height(data, sim::time);
	    compute_pressure(data);
	}
	viscosity::update_viscosity(data);
	compute_heating_cooling_for_CFL(data, sim::time);
	boundary_conditions::copy_initial_values(data);
    }

    // main(): first time step estimate (src/main.cpp:117)
    sim::CalculateTimeStep(data);
    CommunicateBoundariesAll(data);
    CommunicateBoundariesAllInitial(data);

    // metadata
    {
	const std::string path = g_out + "/meta.txt";
	FILE *f = fopen(path.c_str(), "w");
	fprintf(f, "Nr = %u\nNphi = %u\n", Nr, Ns);
	fprintf(f, "G = %.17g\nR = %.17g\nsigma_sb = %.17g\nc = %.17g\n",
		constants::G.get_code_value(), constants::R.get_code_value(),
		constants::sigma.get_code_value(), constants::c.get_code_value());
	fprintf(f, "sigma0 = %.17g\nhydro_center_mass = %.17g\n", parameters::sigma0, hydro_center_mass);
	fprintf(f, "RMIN = %.17g\nRMAX = %.17g\ndphi = %.17g\n", RMIN, RMAX, dphi);
	fprintf(f, "OmegaFrame = %.17g\n", refframe::OmegaFrame);
	fprintf(f, "aspectratio = %.17g\nflaring = %.17g\nalpha = %.17g\n",
		parameters::aspectratio_ref, parameters::flaring_index, parameters::viscous_alpha);
	fprintf(f, "gamma = %.17g\nmu = %.17g\n", parameters::ADIABATICINDEX, parameters::MU);
	fprintf(f, "adiabatic = %d\n", parameters::Adiabatic ? 1 : 0);
	fprintf(f, "cfl = %.17g\ncfl_max_var = %.17g\n", parameters::CFL, parameters::CFL_max_var);
	fprintf(f, "artvisc_factor = %.17g\nartvisc = %d\nartvisc_dissipation = %d\n",
		parameters::artificial_viscosity_factor, (int)parameters::artificial_viscosity,
		parameters::artificial_viscosity_dissipation ? 1 : 0);
	fprintf(f, "thickness_smoothing = %.17g\n", parameters::thickness_smoothing);
	fprintf(f, "sigma_floor = %.17g\n", parameters::sigma_floor);
	fprintf(f, "Tmin = %.17g\nTmax = %.17g\n", parameters::minimum_temperature, parameters::maximum_temperature);
	fprintf(f, "heating_viscous = %d\nheating_viscous_factor = %.17g\n",
		parameters::heating_viscous_enabled ? 1 : 0, parameters::heating_viscous_factor);
	fprintf(f, "cooling_beta_enabled = %d\ncooling_beta = %.17g\ncooling_beta_floor = %d\n",
		parameters::cooling_beta_enabled ? 1 : 0, parameters::cooling_beta,
		parameters::cooling_beta_floor ? 1 : 0);
	fprintf(f, "monitor_timestep = %.17g\n", parameters::monitor_timestep);
	fprintf(f, "radial_viscosity_factor = %.17g\n", parameters::radial_viscosity_factor);
	fprintf(f, "damping_inner_limit = %.17g\ndamping_outer_limit = %.17g\n",
		boundary_conditions::damping_inner_limit, boundary_conditions::damping_outer_limit);
	fprintf(f, "damping_time_factor = %.17g\ndamping_time_radius_outer = %.17g\n",
		boundary_conditions::damping_time_factor, boundary_conditions::damping_time_radius_outer);
	fprintf(f, "damp_inner_scalar = %u\ndamp_inner_vector = %u\n",
		clamp_r_id_to_rmed_grid(get_rmed_id(RMIN * boundary_conditions::damping_inner_limit), false),
		clamp_r_id_to_radii_grid(get_rinf_id(RMIN * boundary_conditions::damping_inner_limit), true));
	fprintf(f, "damp_outer_scalar = %u\ndamp_outer_vector = %u\n",
		clamp_r_id_to_rmed_grid(get_rmed_id(RMAX * boundary_conditions::damping_outer_limit) + 1, false),
		clamp_r_id_to_radii_grid(get_rinf_id(RMAX * boundary_conditions::damping_outer_limit) + 1, true));
	fprintf(f, "n_planets = %u\n", data.get_planetary_system().get_number_of_planets());
	fprintf(f, "damping_enabled = %d\n", boundary_conditions::damping_enabled ? 1 : 0);
	fprintf(f, "cooling_beta_reference = %d\ncooling_beta_ramp_up = %.17g\n",
		parameters::cooling_beta_reference ? 1 : 0, parameters::cooling_beta_ramp_up);
	fprintf(f, "heating_cooling_cfl_limit = %.17g\n", parameters::HEATING_COOLING_CFL_LIMIT);
	fprintf(f, "correct_disk_selfgravity = %d\n", (int)parameters::correct_disk_selfgravity);
	fprintf(f, "compatibility_no_star_smoothing = %d\n", (int)parameters::compatibility_no_star_smoothing);
	fprintf(f, "compatibility_smoothing_planetloc = %d\n", (int)parameters::compatibility_smoothing_planetloc);
	fprintf(f, "integrator = %d\n", (int)parameters::hydro_integrator);
	fclose(f);
    }
    write_raw("geom_Radii.bin", &Radii[0], sizeof(double) * (Nr + 1));
    write_raw("geom_Rinf.bin", &Rinf[0], sizeof(double) * (Nr + 1));
    write_raw("geom_Rsup.bin", &Rsup[0], sizeof(double) * Nr);
    write_raw("geom_Rmed.bin", &Rmed[0], sizeof(double) * (Nr + 1));
    write_raw("geom_Surf.bin", &Surf[0], sizeof(double) * Nr);
    write_raw("geom_InvSurf.bin", &InvSurf[0], sizeof(double) * Nr);
    write_raw("geom_InvRmed.bin", &InvRmed[0], sizeof(double) * (Nr + 1));
    write_raw("geom_InvRinf.bin", &InvRinf[0], sizeof(double) * (Nr + 1));
    write_raw("geom_InvDiffRmed.bin", &InvDiffRmed[0], sizeof(double) * (Nr + 1));
    write_raw("geom_InvDiffRsup.bin", &InvDiffRsup[0], sizeof(double) * Nr);
    write_raw("geom_InvDiffRsupRb.bin", &InvDiffRsupRb[0], sizeof(double) * Nr);
    write_raw("geom_TwoDiffRaSq.bin", &TwoDiffRaSq[0], sizeof(double) * Nr);

    // damping reference values (copied at initialisation)
    dump_grid("init_sigma0", data[t_data::SIGMA0]);
    dump_grid("init_vrad0", data[t_data::V_RADIAL0]);
    dump_grid("init_vazi0", data[t_data::V_AZIMUTHAL0]);
    dump_grid("init_energy0", data[t_data::ENERGY0]);

    if (simrun) {
	options::max_iteration_number = (int)nsteps;
	sim::run(data);
	dump_state(data, "simrun");
	return 0;
    }

    // sim::init (src/simulation.cpp:462-482)
    boundary_conditions::apply_boundary_condition(data, sim::time, 0.0, false);
    refframe::init_corotation(data);
    const double last_dt_before_init = sim::last_dt;
    sim::CalculateTimeStep(data);
    CommunicateBoundaries(&data[t_data::SIGMA], &data[t_data::V_RADIAL],
			  &data[t_data::V_AZIMUTHAL], &data[t_data::ENERGY]);
    {
	FILE *f = fopen((g_out + "/meta.txt").c_str(), "a");
	fprintf(f, "last_dt_before_run_init = %.17g\n", last_dt_before_init);
	fclose(f);
    }
    dump_state(data, "0");

    // sim::run loop (src/simulation.cpp:505-558)
    const bool timing = getenv("REFDUMP_TIME") != nullptr;
    const auto t_loop = std::chrono::steady_clock::now();
    for (unsigned step = 1; step <= nsteps; ++step) {
	double rec[16] = {0};
	rec[0] = sim::time;
	rec[1] = sim::last_dt;
	const double cfl_dt = sim::CalculateTimeStep(data);
	const double time_next_monitor = (sim::N_monitor + 1) * parameters::monitor_timestep;
	const double time_left_till_write = time_next_monitor - sim::time;
	const bool overshoot = cfl_dt > time_left_till_write;
	const double dt_stretch_factor = 0.05;
	const bool almost_there = time_left_till_write < cfl_dt * (1 + dt_stretch_factor);
	const double step_dt = (overshoot || almost_there) ? time_left_till_write : cfl_dt;
	rec[2] = cfl_dt;
	rec[3] = step_dt;
	rec[4] = (double)sim::N_monitor;
	rec[5] = refframe::OmegaFrame;

	std::vector<double> planet_rec;
	if (parameters::hydro_integrator == LEAPFROG_INTEGRATOR) {
	    double ind[6] = {0};
	    leapfrog_step(data, step_dt, step, step <= stage_steps, planet_rec, ind);
	    rec[6] = ind[0];
	    rec[7] = ind[1];
	    rec[10] = ind[4];
	    rec[11] = ind[5];
	} else {
	    euler_step(data, step_dt, step, step <= stage_steps, planet_rec, &rec[6]);
	}
	rec[8] = sim::time;
	const bool towrite = std::fabs(time_next_monitor - sim::time) < 1e-6 * cfl_dt;
	if (towrite) {
	    sim::N_monitor++;
	}
	rec[9] = towrite ? 1.0 : 0.0;
	if (timing) continue;
	append_raw("steps.bin", rec, sizeof(rec));
	append_raw("planets.bin", planet_rec.data(), sizeof(double) * planet_rec.size());
	dump_disk_accel(data);
	dump_totals(data);
	if (checkpoints.count(step)) {
	    dump_state(data, std::to_string(step));
	}
    }
    if (timing) {
	const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_loop).count();
	FILE *f = fopen((g_out + "/meta.txt").c_str(), "a");
	fprintf(f, "loop_seconds = %.9g\n", sec);
	fclose(f);
    }
    if (getenv("REFDUMP_SNAPSHOT")) {
	// the reference's own snapshot writers on the final state
	// (t_polargrid::write2D, polargrid.cpp:135-185; output::write_misc,
	// output.cpp:494-525) with snapshot number 7
	output::snapshot_dir = g_out + "/snapshot";
	std::string cmd = "mkdir -p " + output::snapshot_dir;
	if (system(cmd.c_str()) != 0) return 3;
	const t_data::t_polargrid_type grids[] = {t_data::SIGMA, t_data::V_RADIAL, t_data::V_AZIMUTHAL,
						  t_data::ENERGY, t_data::QPLUS, t_data::QMINUS,
						  t_data::TEMPERATURE, t_data::PRESSURE};
	for (auto g : grids) {
	    data[g].write2D(output::snapshot_dir + "/" + std::string(data[g].get_name()) + ".dat");
	}
	sim::N_snapshot = 7;
	output::write_misc();
    }
    return 0;
}
