// Device-side data structures shared by the gfx950 kernels and the host glue.
//
// One disk = one KParams block passed by value to every kernel: grid sizes,
// physical scalars, pointers to the per-ring metric arrays and to every field
// in HBM (and, for split storage, to the per-ring backgrounds).  Quantities that change from step to step (dt, time, planets, the
// CFL minimum) live in a DevScalars block in device memory so that a captured
// graph of the step can be replayed without host involvement.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "../../include/fcpt.h"

namespace fcpt {

constexpr int kMaxPlanets = FCPT_MAX_PLANETS;
constexpr unsigned kCflMaxBlocks = 4096;  // slots of the block-minimum array
// rows of KParams::ring_tab
enum { TAB_SRC_INVDXTHETA = 0, TAB_INV_DXTHETA, TAB_INV_OMEGA_K, TAB_OMEGA_K, TAB_VISC_AREA, kRingTabs };

struct DevPlanet {
    double x, y, mass, r_cubic;
    double smooth;  // fixed potential smoothing length, or < 0: ThicknessSmoothing * H of the cell
};

// per-body potential smoothing of one step, formed on the host (compute_smoothing)
struct BodySmoothing {
    double s[kMaxPlanets];
};

// Step-varying scalars, resident in device memory.
struct DevScalars {
    double dt;            // step size used by the current step's kernels
    double time;          // simulation time at the start of the step
    double last_dt;       // sim::last_dt (CFLmaxVar limiter)
    double cfl_dt;        // result of the last CFL reduction
    int n_monitor;
    int error_flags;      // bit 0: non-finite or non-positive dt
    int n_planets;
    unsigned step_count;  // CFL reductions so far: part of the split storage's rounding key
    double indirect_x, indirect_y;
    DevPlanet planet[kMaxPlanets];
    // LeapFrog: the full step's dt and start time while the kicks run with
    // dt / 2, the second kick's bodies (has_mid) and the first kick's, kept
    // while the second kick uses the slots above
    double dt_step, t_step;
    int has_mid, n_planets_mid, n_planets_first, pad2_;
    double indirect_mid_x, indirect_mid_y, indirect_first_x, indirect_first_y;
    DevPlanet planet_mid[kMaxPlanets];
    DevPlanet planet_first[kMaxPlanets];
};

struct KParams {
    int nr, ns;
    double dphi, invdphi;
    double rmin, rmax;
    // physics
    int adiabatic;
    int integrator;      // FCPT_INTEGRATOR_*
    int artvisc;
    int artvisc_dissipation;
    int heating_viscous;
    int cooling_beta_enabled;
    int cooling_beta_ref;
    int bc_inner, bc_outer;
    int damping;
    int damp_ids[4];
    double G, Rgas, sigma_sb, c_light, mstar, mu, gamma;
    double aspect_ratio, flaring, alpha, thickness_smoothing;
    double cfl, cfl_max_var, heating_cooling_cfl_limit;
    double sigma0, sigma_floor, t_min, t_max;
    double artvisc_factor, heating_viscous_factor;
    double cooling_beta, cooling_beta_ramp_up, radial_viscosity_factor;
    double damping_inner_limit, damping_outer_limit, damping_time_factor,
        damping_time_radius_outer;
    double omega_frame, monitor_timestep;
    // per-ring metric arrays (length nr + 2)
    const double *Rinf, *Rsup, *Rmed, *Surf, *InvSurf, *InvRmed, *InvRinf, *InvDiffRmed,
        *InvDiffRsup, *InvDiffRsupRb;
    // per-sector tables: cos / sin of dphi * j (cell-centre coordinates)
    const double *cosphi, *sinphi;
    // storage: 0 = FP64 fields; 1 = FP32 deviations from FP64 per-ring backgrounds
    int split;
    int split_rounding;  // split storage: 0 rounds deviations to nearest, 1 stochastically
    int radial_tile;     // rings per thread of the radial sweep (0: size-based choice)
    int radial_waves;    // register budget of the radial sweep: wavefronts per SIMD (2 or 3)
    int visc_tile;       // rings per thread of the fused viscous pass (0: size-based choice)
    int kick_tile;       // rings per thread of the kick walk (0: size-based choice, 1: off)
    // the fused Euler step of an adiabatic disk (either storage): the sound
    // speed, viscosity and temperature are formed from E and Sigma where the
    // step needs them instead of stored (the kick walk stores nu only); the
    // host stores them all (k_derived) before anything else reads the fields
    int lazy_derived;
    // the fused step's walks of an adiabatic disk (split storage): the kick walk
    // forms the pressure from E and writes the updated E to the partner buffer
    // ealt, which the viscous walk reads (its SubStep3 writes E back); the end
    // of the step does not store the pressure
    int e_partner;
    int cfl_gx, cfl_gy;  // grid of the CFL cell kernel (its block-minimum slots)
    // fields in HBM (indexed by fcpt_field): double, or float when split
    void *f[FCPT_NUM_FIELDS];
    // split storage: per-ring backgrounds of every field (nr + 2 values each;
    // null in FP64 mode); the potential's is rewritten by the potential kernel
    double *bg[FCPT_NUM_FIELDS];
    // transport work arrays: the momenta between the sweeps (mom_b) have the
    // fields' element type; the azimuthal sweep's momenta (mom, read by the
    // end of the step) and the density and energy between the sweeps are
    // FP64 in both storage modes
    void *mom[4];        // radial momentum +/-, angular momentum +/- (FP64)
    void *mom_b[4];      // the radial sweep's momenta
    void *sig_b, *en_b;  // radial-sweep outputs
    void *sig_c, *en_c;  // azimuthal-sweep outputs
    void *valt[2];       // partners of v_r and v_phi for the fused step kernels
    void *ealt;          // partner of E (adiabatic disks in split storage; e_partner)
    double *bgm[4];      // backgrounds of the four momenta (split storage)
    double *iso_cs, *iso_h, *iso_nu;  // locally isothermal ring profiles of c_s, H, nu
    // ring-uniform quotients of the source terms, the artificial viscosity,
    // the viscous update and the derived fields (k_ring_tables), kRingTabs
    // rows of nr + 2 values
    double *ring_tab;
    double *vmean;       // per-ring mean azimuthal velocity
    double *vuni;        // per-ring uniform residual velocity (FARGO)
    int *nshift;         // per-ring integer shift
    double *ring_min;    // CFL minima: per block (kCflMaxBlocks slots) or per ring (nr)
    DevScalars *s;
};

}  // namespace fcpt
