// C++ host interface over the C ABI (see fcpt_host.hpp).
#include "fcpt_host.hpp"

#include <cmath>

namespace fcpt {

void check(fcpt_ctx *ctx, int status)
{
    if (status != FCPT_OK) throw Error(status, ctx ? fcpt_last_error(ctx) : "fcpt: invalid argument");
}

Data::Data(const fcpt_disk_desc &desc)
    : desc_(desc), nr_(desc.nr), nphi_(desc.nphi), adiabatic_(desc.adiabatic != 0)
{
    desc_.radii = nullptr;
    check(nullptr, fcpt_create(&desc, &ctx_));
}

Data::~Data() { fcpt_destroy(ctx_); }

void Data::upload(const std::vector<double> &sigma, const std::vector<double> &vrad,
                  const std::vector<double> &vazi, const std::vector<double> &energy)
{
    const size_t n = (size_t)nr_ * nphi_;
    if (sigma.size() != n || vrad.size() != n + nphi_ || vazi.size() != n ||
        (adiabatic_ && energy.size() != n))
        throw Error(FCPT_ERR_ARG, "fcpt::Data::upload: array sizes do not match the grid");
    check(ctx_, fcpt_upload_state(ctx_, sigma.data(), vrad.data(), vazi.data(),
                                  adiabatic_ ? energy.data() : nullptr));
}

void Data::upload_reference(const std::vector<double> &sigma0, const std::vector<double> &vrad0,
                            const std::vector<double> &vazi0, const std::vector<double> &energy0)
{
    check(ctx_, fcpt_upload_reference(ctx_, sigma0.data(), vrad0.data(), vazi0.data(),
                                      energy0.empty() ? nullptr : energy0.data()));
}

std::vector<double> Data::field(int id) const
{
    const bool vec = id == FCPT_VRAD || id == FCPT_VRAD0 || id == FCPT_TAU_R_PHI;
    std::vector<double> out((size_t)(vec ? nr_ + 1 : nr_) * nphi_);
    check(ctx_, fcpt_download_field(ctx_, id, out.data()));
    return out;
}

void Data::set_bodies(const std::vector<Body> &bodies, double indirect_x, double indirect_y)
{
    if (bodies.size() > FCPT_MAX_PLANETS) throw Error(FCPT_ERR_ARG, "too many bodies");
    fcpt_step_params p{};
    p.n_planets = (int32_t)bodies.size();
    for (size_t k = 0; k < bodies.size(); ++k)
        p.planet[k] = {bodies[k].x, bodies[k].y, bodies[k].mass, bodies[k].r_cubic};
    p.indirect_x = indirect_x;
    p.indirect_y = indirect_y;
    check(ctx_, fcpt_set_step_params(ctx_, &p));
}

namespace cfl {
double condition_cfl(Data &data, double)
{
    double dt = 0.0;
    check(data.ctx(), fcpt_condition_cfl(data.ctx(), &dt));
    return dt;
}
}  // namespace cfl

namespace sim {
double time = 0.0;
double last_dt = 0.0;
unsigned int N_monitor = 0;
unsigned long N_hydro_iter = 0;

static void push_clock(Data &data)
{
    check(data.ctx(), fcpt_set_time(data.ctx(), time, last_dt, (int32_t)N_monitor));
}

static void pull_clock(Data &data)
{
    int32_t nm = 0;
    check(data.ctx(), fcpt_get_time(data.ctx(), &time, &last_dt, &nm));
    N_monitor = (unsigned)nm;
}

double CalculateTimeStep(Data &data)
{
    push_clock(data);
    double dt = 0.0;
    check(data.ctx(), fcpt_calculate_timestep(data.ctx(), &dt));
    last_dt = dt;
    return dt;
}

void init(Data &data)
{
    boundary_conditions::apply_boundary_condition(data, time, 0.0, false);
    CalculateTimeStep(data);
}

// step_Euler, gas part (simulation.cpp:148-267), one call per reference function
static void step_Euler(Data &data, double dt)
{
    CalculateNbodyPotential(data, time);
    update_with_sourceterms(data, dt);
    art_visc::update_with_artificial_viscosity(data, dt);
    recalculate_viscosity(data, time);
    viscosity::compute_viscous_stress_tensor(data);
    viscosity::update_velocities_with_viscosity(data, dt);
    if (data.adiabatic()) SubStep3(data, time, dt);
    boundary_conditions::apply_boundary_condition(data, time, 0.0, false);
    Transport(data, dt);
    time += dt;
    N_hydro_iter = N_hydro_iter + 1;
    boundary_conditions::apply_boundary_condition(data, time, dt, true);
    recalculate_derived_disk_quantities(data, time);
}

void run(Data &data, unsigned long n_steps, const NbodyHook &nbody)
{
    // the loop body of sim::run (simulation.cpp:515-556) with dt, the monitor
    // clip and the time kept on the device
    for (unsigned long k = 0; k < n_steps; ++k) {
        if (nbody) nbody(data, time);
        push_clock(data);
        double dt_step = 0.0;
        check(data.ctx(), fcpt_run_step(data.ctx(), &dt_step));
        pull_clock(data);
        N_hydro_iter = N_hydro_iter + 1;
    }
}

void run_stepwise(Data &data, unsigned long n_steps, const NbodyHook &nbody)
{
    // the same loop with one call per reference function and the clip on the host
    for (unsigned long k = 0; k < n_steps; ++k) {
        if (nbody) nbody(data, time);
        const double cfl_dt = CalculateTimeStep(data);
        const double time_next_monitor = (N_monitor + 1) * data.desc().monitor_timestep;
        const double time_left_till_write = time_next_monitor - time;
        const bool overshoot = cfl_dt > time_left_till_write;
        const double dt_stretch_factor = 0.05;
        const bool almost_there = time_left_till_write < cfl_dt * (1 + dt_stretch_factor);
        const double step_dt = (overshoot || almost_there) ? time_left_till_write : cfl_dt;
        step_Euler(data, step_dt);
        if (std::fabs(time_next_monitor - time) < 1e-6 * cfl_dt) N_monitor++;
    }
    push_clock(data);
}
}  // namespace sim

void CalculateNbodyPotential(Data &data, double) { check(data.ctx(), fcpt_nbody_potential(data.ctx())); }

void update_with_sourceterms(Data &data, double dt)
{
    check(data.ctx(), fcpt_update_with_sourceterms(data.ctx(), dt));
}

void recalculate_viscosity(Data &data, double) { check(data.ctx(), fcpt_recalculate_viscosity(data.ctx())); }

void SubStep3(Data &data, double current_time, double dt)
{
    check(data.ctx(), fcpt_substep3(data.ctx(), current_time, dt));
}

void recalculate_derived_disk_quantities(Data &data, double)
{
    check(data.ctx(), fcpt_recalculate_derived(data.ctx()));
}

void Transport(Data &data, double dt) { check(data.ctx(), fcpt_transport(data.ctx(), dt)); }

namespace art_visc {
void update_with_artificial_viscosity(Data &data, double dt)
{
    check(data.ctx(), fcpt_artificial_viscosity(data.ctx(), dt));
}
}  // namespace art_visc

namespace viscosity {
void compute_viscous_stress_tensor(Data &data)
{
    check(data.ctx(), fcpt_viscous_stress_tensor(data.ctx()));
}
void update_velocities_with_viscosity(Data &data, double dt)
{
    check(data.ctx(), fcpt_update_velocities_with_viscosity(data.ctx(), dt));
}
}  // namespace viscosity

namespace boundary_conditions {
void apply_boundary_condition(Data &data, double, double dt, bool final)
{
    check(data.ctx(), fcpt_apply_boundary_condition(data.ctx(), dt, final ? 1 : 0));
}
}  // namespace boundary_conditions

}  // namespace fcpt
