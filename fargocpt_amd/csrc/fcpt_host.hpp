// C++ host interface over the C ABI, mirroring the reference's hydro-step
// functions: same names, same argument meaning, same call order in
// sim::step / sim::run.  Where the reference takes t_data& this layer takes
// fcpt::Data&, which owns one GPU-resident disk; fatal conditions that make
// the reference die() throw fcpt::Error here.
#pragma once

#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/fcpt.h"

namespace fcpt {

struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string &m) : std::runtime_error(m), status(s) {}
};

// Host-side inputs the reference's N-body module publishes each step
// (g_xpl/g_ypl/g_mpl/g_cubic_smoothing_radius and refframe::IndirectTerm).
struct Body {
    double x, y, mass, r_cubic;
};

class Data {
  public:
    explicit Data(const fcpt_disk_desc &desc);
    ~Data();
    Data(const Data &) = delete;
    Data &operator=(const Data &) = delete;

    fcpt_ctx *ctx() const { return ctx_; }
    int nr() const { return nr_; }
    int nphi() const { return nphi_; }
    bool adiabatic() const { return adiabatic_; }
    const fcpt_disk_desc &desc() const { return desc_; }

    // state transfer in the reference layout
    void upload(const std::vector<double> &sigma, const std::vector<double> &vrad,
                const std::vector<double> &vazi, const std::vector<double> &energy);
    void upload_reference(const std::vector<double> &sigma0, const std::vector<double> &vrad0,
                          const std::vector<double> &vazi0, const std::vector<double> &energy0);
    std::vector<double> field(int id) const;

    // the N-body inputs of the next step
    void set_bodies(const std::vector<Body> &bodies, double indirect_x, double indirect_y);

  private:
    fcpt_disk_desc desc_{};
    fcpt_ctx *ctx_ = nullptr;
    int nr_ = 0, nphi_ = 0;
    bool adiabatic_ = false;
};

void check(fcpt_ctx *ctx, int status);

namespace cfl {
// cfl::condition_cfl (cfl.h:8)
double condition_cfl(Data &data, double dt_global_input = 0.0);
}  // namespace cfl

namespace sim {
extern double time;
extern double last_dt;
extern unsigned int N_monitor;
extern unsigned long N_hydro_iter;
// sim::CalculateTimeStep (simulation.h:19)
double CalculateTimeStep(Data &data);
// Host-side N-body hook: publishes the bodies and indirect term of the step
// that starts at `time` (Data::set_bodies).  Empty: keep the last ones.
using NbodyHook = std::function<void(Data &, double time)>;
// sim::init and the loop of sim::run (simulation.cpp:462-558), n_steps iterations;
// run() keeps dt on the device (fcpt_run_step), run_stepwise() makes one call
// per reference function like step_Euler
void init(Data &data);
void run(Data &data, unsigned long n_steps, const NbodyHook &nbody = {});
void run_stepwise(Data &data, unsigned long n_steps, const NbodyHook &nbody = {});
}  // namespace sim

// Pframeforce.h
void CalculateNbodyPotential(Data &data, double current_time);
// SourceEuler.h:11-18
void update_with_sourceterms(Data &data, double dt);
void recalculate_viscosity(Data &data, double current_time);
void SubStep3(Data &data, double current_time, double dt);
void recalculate_derived_disk_quantities(Data &data, double current_time);
// TransportEuler.h:5-6 (the four PolarGrid* of the reference are Data's own fields)
void Transport(Data &data, double dt);

namespace art_visc {
void update_with_artificial_viscosity(Data &data, double dt);
}
namespace viscosity {
void compute_viscous_stress_tensor(Data &data);
void update_velocities_with_viscosity(Data &data, double dt);
}  // namespace viscosity
namespace boundary_conditions {
void apply_boundary_condition(Data &data, double current_time, double dt, bool final);
}

}  // namespace fcpt
